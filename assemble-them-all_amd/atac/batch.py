"""Python front end of the contact library: grids, point sets and batches of
force objects held on the GPU, evaluated through the C ABI. torch provides the
device buffers and the stream; the arithmetic runs in the HIP kernels."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from ._lib import check, lib


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def device_init(device: int = 0) -> int:
    """Select the device for this thread; returns its compute-unit count."""
    n = C.c_int32(0)
    check(lib().atac_device_init(device, C.byref(n)))
    return int(n.value)


class DeviceGrid:
    """FP32 SDF grid in HBM (BodySDFObj::_SDF with its box and spacing)."""

    def __init__(self, values, min_box, max_box, dxyz):
        v = np.ascontiguousarray(values, dtype=np.float32)
        if v.ndim != 3:
            raise ValueError("values must have shape (nk, nj, ni)")
        nk, nj, ni = v.shape
        self.shape = (ni, nj, nk)
        self.min_box = np.asarray(min_box, dtype=np.float64).copy()
        self.max_box = np.asarray(max_box, dtype=np.float64).copy()
        self.dxyz = np.asarray(dxyz, dtype=np.float64).copy()
        h = C.c_void_p()
        check(lib().atac_grid_create(v.ctypes.data_as(C.c_void_p), ni, nj, nk, L.dptr(self.min_box),
                                     L.dptr(self.max_box), L.dptr(self.dxyz), C.byref(h)))
        self.handle = h

    @classmethod
    def from_grid(cls, g):
        return cls(g.values, g.min_box, g.max_box, g.dxyz)

    def close(self):
        if self.handle:
            lib().atac_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def distance(self, xw, E2=None, col_th=0.0):
        """BodySDFObj::distance / SDFMesh::distance of world points (torch or numpy, n x 3)."""
        E = np.eye(4) if E2 is None else np.ascontiguousarray(E2, dtype=np.float64)
        x = torch.as_tensor(xw, dtype=torch.float64).reshape(-1, 3).contiguous().cuda()
        out = torch.empty(x.shape[0], dtype=torch.float64, device=x.device)
        check(lib().atac_sdf_distance(self.handle, L.dptr(E), float(col_th), _ptr(x), x.shape[0], _ptr(out), _stream()))
        return out

    def collision(self, xw, xw_dot, E2, phi2, col_th, derivs=False, precise=False):
        """BodySDFObj::collision for n points: values (n, 11) and optionally the 135 derivative values."""
        E = np.ascontiguousarray(E2, dtype=np.float64)
        ph = np.ascontiguousarray(phi2, dtype=np.float64)
        x = torch.as_tensor(xw, dtype=torch.float64).reshape(-1, 3).contiguous().cuda()
        xd = torch.as_tensor(xw_dot, dtype=torch.float64).reshape(-1, 3).contiguous().cuda()
        vals = torch.empty((x.shape[0], 11), dtype=torch.float64, device=x.device)
        der = torch.empty((x.shape[0], 135), dtype=torch.float64, device=x.device) if derivs else None
        check(lib().atac_sdf_collision(self.handle, L.dptr(E), L.dptr(ph), float(col_th),
                                       L.EVAL_PRECISE if precise else 0, _ptr(x), _ptr(xd), x.shape[0], _ptr(vals),
                                       _ptr(der), _stream()))
        return vals, der


class DevicePoints:
    """Contact-point sets (Body::_contact_points), Morton-sorted SoA FP32 in HBM."""

    def __init__(self, xyz, seg_offsets=None, sort_cell=0.0):
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.n = len(x)
        if seg_offsets is None:
            self.seg = np.array([0, self.n], dtype=np.int64)
            segp, nseg = None, 1
        else:
            self.seg = np.ascontiguousarray(seg_offsets, dtype=np.int64)
            segp, nseg = self.seg.ctypes.data_as(C.c_void_p), len(self.seg) - 1
        h = C.c_void_p()
        check(lib().atac_points_create(x.ctypes.data_as(C.c_void_p), self.n, segp, nseg, float(sort_cell), C.byref(h)))
        self.handle = h

    @property
    def n_segments(self):
        return len(self.seg) - 1

    def permutation(self, segment=0):
        m = int(self.seg[segment + 1] - self.seg[segment])
        out = np.empty(m, dtype=np.int64)
        check(lib().atac_points_permutation(self.handle, segment, out.ctypes.data_as(C.c_void_p), m))
        return out

    def close(self):
        if self.handle:
            lib().atac_points_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_instance(grid, points, segment, E1, phi1, E2, phi2, col_th, kn, kt, mu, damping, scale=1.0,
                  contact_scale=1.0) -> L.Instance:
    s = L.Instance()
    s.grid, s.points, s.segment = int(grid), int(points), int(segment)
    s.E1[:] = [float(v) for v in np.asarray(E1, dtype=np.float64).ravel()]
    s.phi1[:] = [float(v) for v in np.asarray(phi1, dtype=np.float64).ravel()]
    s.E2[:] = [float(v) for v in np.asarray(E2, dtype=np.float64).ravel()]
    s.phi2[:] = [float(v) for v in np.asarray(phi2, dtype=np.float64).ravel()]
    s.col_th, s.kn, s.kt, s.mu, s.damping = float(col_th), float(kn), float(kt), float(mu), float(damping)
    s.scale, s.contact_scale = float(scale), float(contact_scale)
    return s


class ContactBatch:
    """Many force objects (ForceGeneralSDFContact) evaluated in one persistent launch."""

    def __init__(self, grids, point_sets, instances, tuning=None):
        self.grids = list(grids)
        self.point_sets = list(point_sets)
        self.n = len(instances)
        arr = (L.Instance * self.n)(*instances)
        self._inst = arr
        garr = (C.c_void_p * len(self.grids))(*[g.handle for g in self.grids])
        parr = (C.c_void_p * len(self.point_sets))(*[p.handle for p in self.point_sets])
        tu = self._tuning(tuning)
        h = C.c_void_p()
        check(lib().atac_batch_create(garr, len(self.grids), parr, len(self.point_sets), arr, self.n,
                                      C.byref(tu) if tu is not None else None, C.byref(h)))
        self.handle = h

    @property
    def instances(self):
        """The current Instance records (a ctypes array; set_instances accepts it back)."""
        return self._inst

    @staticmethod
    def _tuning(tuning):
        if tuning is None:
            return None
        if isinstance(tuning, L.Tuning):
            return tuning
        t = L.Tuning()
        t.tile_points = int(tuning.get("tile_points", 0))
        t.detect_block = int(tuning.get("detect_block", 0))
        t.blocks_per_cu = int(tuning.get("blocks_per_cu", 0))
        t.order = int(tuning.get("order", 0))
        t.points_per_lane = int(tuning.get("points_per_lane", 0))
        t.max_active = int(tuning.get("max_active", 0))
        t.group_cull = int(tuning.get("group_cull", 0))
        return t

    def close(self):
        if getattr(self, "handle", None):
            lib().atac_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n = C.c_int64()
        nt = C.c_int32()
        pb = C.c_int32()
        tu = L.Tuning()
        check(lib().atac_batch_info(self.handle, C.byref(n), C.byref(nt), C.byref(pb), C.byref(tu)))
        return {"n_points": int(n.value), "n_tiles": int(nt.value), "persistent_blocks": int(pb.value),
                "tile_points": tu.tile_points, "detect_block": tu.detect_block, "blocks_per_cu": tu.blocks_per_cu,
                "order": tu.order, "points_per_lane": tu.points_per_lane, "max_active": tu.max_active,
                "group_cull": tu.group_cull}

    def retune(self, tuning):
        self.generation += 1
        tu = self._tuning(tuning)
        check(lib().atac_batch_retune(self.handle, C.byref(tu)))

    def set_instances(self, instances):
        """Instances as a sequence of Instance records (or a ctypes array of them)."""
        arr = instances if isinstance(instances, C.Array) else (L.Instance * self.n)(*instances)
        self._inst = arr
        self.generation += 1   # the adjoint records of the last evaluation belong to the old instances
        check(lib().atac_batch_set_instances(self.handle, arr, self.n, _stream()))

    def alloc_outputs(self, deriv=False, p6=False, device="cuda"):
        f64 = dict(dtype=torch.float64, device=device)
        out = {"fm": torch.empty((self.n, 12), **f64),
               "n_active": torch.empty(self.n, dtype=torch.int32, device=device)}
        if p6 or deriv:
            out["dfm_dp6"] = torch.empty((self.n, 12), **f64)
        if deriv:
            out["Km"] = torch.empty((self.n, 12, 12), **f64)
            out["Dm"] = torch.empty((self.n, 12, 12), **f64)
        return out

    @staticmethod
    def _outputs_struct(out):
        o = L.Outputs()
        o.fm = out["fm"].data_ptr() if "fm" in out else None
        o.Km = out["Km"].data_ptr() if "Km" in out else None
        o.Dm = out["Dm"].data_ptr() if "Dm" in out else None
        o.dfm_dp6 = out["dfm_dp6"].data_ptr() if "dfm_dp6" in out else None
        o.n_active = out["n_active"].data_ptr() if "n_active" in out else None
        return o

    @staticmethod
    def flags(deriv=False, p6=False, checkpoint=False, precise=False, all_active=False):
        return ((L.EVAL_DERIV if deriv else 0) | (L.EVAL_P6 if p6 else 0) | (L.EVAL_CHECKPOINT if checkpoint else 0)
                | (L.EVAL_PRECISE if precise else 0) | (L.EVAL_ALL_ACTIVE if all_active else 0))

    generation = 0   # evaluations (and re-plans) so far; the adjoint differentiates the last one

    def eval(self, deriv=False, p6=False, checkpoint=False, precise=False, all_active=False, out=None):
        out = out if out is not None else self.alloc_outputs(deriv, p6)
        o = self._outputs_struct(out)
        self.generation += 1
        check(lib().atac_batch_eval(self.handle, self.flags(deriv, p6, checkpoint, precise, all_active), C.byref(o),
                                    _stream()))
        return out

    def eval_vjp(self, g_fm, precise=False, all_active=False, out=None, grads=None):
        """eval() then vjp(g_fm) for a cotangent known beforehand, in one fused pass over the
        contacts (wrench only); it keeps no records, so vjp() needs a new eval() after it."""
        g_fm = g_fm.to(dtype=torch.float64).contiguous()
        out = out if out is not None else self.alloc_outputs(False, False)
        grads = grads if grads is not None else self.alloc_gradients(g_fm.device)
        o, gs = self._outputs_struct(out), self._grad_struct(grads)
        self.generation += 1
        check(lib().atac_batch_eval_vjp(self.handle, self.flags(False, False, False, precise, all_active), C.byref(o),
                                        _ptr(g_fm), C.byref(gs), _stream()))
        return out, grads

    def eval_plain(self, deriv=False, p6=False, precise=False, out=None):
        out = out if out is not None else self.alloc_outputs(deriv, p6)
        o = self._outputs_struct(out)
        check(lib().atac_eval_plain(self.handle, self.flags(deriv, p6, False, precise), C.byref(o), _stream()))
        return out

    def eval_plain_vjp(self, g_fm, precise=False, out=None, grads=None):
        """The plain per-instance path with the outputs of eval() + vjp(g_fm)."""
        g_fm = g_fm.to(dtype=torch.float64).contiguous()
        out = out if out is not None else self.alloc_outputs()
        grads = grads if grads is not None else self.alloc_gradients(g_fm.device)
        o = self._outputs_struct(out)
        gs = self._grad_struct(grads)
        check(lib().atac_eval_plain_vjp(self.handle, self.flags(precise=precise), _ptr(g_fm), C.byref(o), C.byref(gs),
                                        _stream()))
        return out, grads

    def alloc_gradients(self, device="cuda"):
        f64 = dict(dtype=torch.float64, device=device)
        return {"g_q1": torch.empty((self.n, 6), **f64), "g_q2": torch.empty((self.n, 6), **f64),
                "g_phi1": torch.empty((self.n, 6), **f64), "g_phi2": torch.empty((self.n, 6), **f64),
                "g_kn": torch.empty(self.n, **f64), "g_kt": torch.empty(self.n, **f64)}

    @staticmethod
    def _grad_struct(g):
        s = L.Gradients()
        for k in ("g_q1", "g_q2", "g_phi1", "g_phi2", "g_kn", "g_kt"):
            setattr(s, k, g[k].data_ptr() if k in g else None)
        return s

    def vjp(self, g_fm, out=None):
        g_fm = g_fm.to(dtype=torch.float64).contiguous()
        out = out if out is not None else self.alloc_gradients(g_fm.device)
        s = self._grad_struct(out)
        check(lib().atac_batch_vjp(self.handle, _ptr(g_fm), C.byref(s), _stream()))
        return out

    def graph_capture(self, out, deriv=False, p6=False, checkpoint=False, precise=False, g_fm=None, grads=None,
                      fused=False):
        o = self._outputs_struct(out)
        gs = self._grad_struct(grads) if grads is not None else None
        fl = self.flags(deriv, p6, checkpoint, precise) | (L.EVAL_FUSED_ADJOINT if fused else 0)
        check(lib().atac_batch_graph_capture(self.handle, fl, C.byref(o),
                                             _ptr(g_fm) if g_fm is not None else None,
                                             C.byref(gs) if gs is not None else None))

    def graph_launch(self):
        self.generation += 1
        check(lib().atac_batch_graph_launch(self.handle, _stream()))

    def active_ids(self, i):
        cnt = C.c_int64()
        check(lib().atac_batch_active_ids(self.handle, i, None, 0, C.byref(cnt)))
        ids = np.empty(cnt.value, dtype=np.int64)
        check(lib().atac_batch_active_ids(self.handle, i, ids.ctypes.data_as(C.c_void_p), len(ids), C.byref(cnt)))
        return ids

    STAGES = ("detect", "slow", "contact", "reduce", "vjp", "vjp_reduce")

    def timing(self, enable=True):
        """Read (and clear) per-stage kernel times recorded with HIP events, then
        switch recording on or off. Returns {stage: (total_ms, launches)}."""
        ms = np.zeros(len(self.STAGES))
        n = np.zeros(len(self.STAGES), dtype=np.int32)
        check(lib().atac_batch_timing(self.handle, 1 if enable else 0, ms.ctypes.data_as(C.c_void_p),
                                      n.ctypes.data_as(C.c_void_p)))
        return {s: (float(ms[k]), int(n[k])) for k, s in enumerate(self.STAGES)}

    def workspace(self):
        """The record pool: {capacity, used (contacts of the last evaluation), overflowed}. Synchronous."""
        cap, used, ovf = C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().atac_batch_workspace(self.handle, C.byref(cap), C.byref(used), C.byref(ovf)))
        return {"capacity": int(cap.value), "used": int(used.value), "overflowed": bool(ovf.value)}

    def live_points(self):
        """(points the last evaluation's detection read after group culling, points of the batch).
        Synchronous."""
        live, n = C.c_int64(), C.c_int64()
        check(lib().atac_batch_live_points(self.handle, C.byref(live), C.byref(n)))
        return int(live.value), int(n.value)

    def touched_values(self):
        c = C.c_int64()
        check(lib().atac_batch_touched_values(self.handle, C.byref(c)))
        return int(c.value)


def batch_from_workload(w, tuning=None, sort_cell=0.0, instances=None):
    """Device grids, one segmented point set and the batch for a workloads.Workload."""
    grids = [DeviceGrid.from_grid(g) for g in w.grids]
    sizes = np.array([len(p) for p in w.point_sets], dtype=np.int64)
    seg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    allp = np.concatenate(w.point_sets, axis=0) if len(w.point_sets) else np.zeros((0, 3))
    pts = DevicePoints(allp, seg, sort_cell)
    inst = [make_instance(i.grid, 0, i.points, i.E1, i.phi1, i.E2, i.phi2, i.col_th, i.kn, i.kt, i.mu, i.damping,
                          i.scale, i.contact_scale) for i in (instances if instances is not None else w.instances)]
    return ContactBatch(grids, [pts], inst, tuning)

"""The fused wrench-and-adjoint pass (atac_batch_eval_vjp, and its captured graph) against
the evaluation followed by the adjoint (atac_batch_eval + atac_batch_vjp) on the same
inputs: the same wrench, active counts and gradients, in both arithmetics."""
import numpy as np
import pytest
import torch

from atac import workloads as W
from atac.batch import batch_from_workload
from atac._lib import AtacError

pytestmark = pytest.mark.gpu

KEYS = ["g_q1", "g_q2", "g_phi1", "g_phi2", "g_kn", "g_kt"]


def _np(d):
    return {k: v.cpu().numpy() for k, v in d.items()}


def _worst(ref, got):
    # per output, relative to its largest entry over the batch
    return {k: float(np.abs(got[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-300)) for k in ref}


# The two paths run the same per-contact arithmetic and sum in the same order (per lane
# over the records 64 apart, then over the lanes, then over the items in tile order); they
# are separate kernels, so the compiler may contract a product into an FMA in one and not
# the other: agreement to 1e-6 of each output's largest entry (FP32) and 1e-12 (FP64).
@pytest.mark.parametrize("precise,tol", [(False, 1e-6), (True, 1e-12)])
def test_fused_equals_eval_then_vjp(gpu, precise, tol):
    w = W.peg_in_hole(n_poses=32)
    b = batch_from_workload(w)
    g = torch.tensor(np.random.default_rng(3).normal(size=(b.n, 12)), device="cuda")
    print(f"[case] peg_in_hole 32 poses, precise={precise}", flush=True)
    out = _np(b.eval(checkpoint=True, precise=precise))
    grads = _np(b.vjp(g))
    fo, fg = b.eval_vjp(g, precise=precise)
    fo, fg = _np(fo), _np(fg)
    assert np.array_equal(fo["n_active"], out["n_active"]) and out["n_active"].sum() > 0
    worst = _worst({"fm": out["fm"], **grads}, {"fm": fo["fm"], **fg})
    print(f"  worst relative difference {max(worst.values()):.3g} ({max(worst, key=worst.get)})", flush=True)
    for k, e in worst.items():
        assert e <= tol, k
    # no records kept: the adjoint needs a new evaluation
    with pytest.raises(AtacError):
        b.vjp(g)
    b.eval(checkpoint=True, precise=precise)
    b.vjp(g)


def test_fused_graph_replays_equal_eager(gpu):
    w = W.batched(n_inst=256, n_min=500, n_max=20_000)
    b = batch_from_workload(w)
    g = torch.tensor(np.random.default_rng(4).normal(size=(b.n, 12)), device="cuda")
    print(f"[case] batched 256 instances, {w.n_queries} queries: fused graph x3 vs eager", flush=True)
    eo, eg = b.eval_vjp(g)
    eo, eg = _np(eo), _np(eg)
    fo, fg = b.alloc_outputs(), b.alloc_gradients()
    b.graph_capture(fo, g_fm=g, grads=fg, fused=True)
    for _ in range(3):
        b.graph_launch()
        torch.cuda.synchronize()
        ro, rg = _np(fo), _np(fg)
        assert np.array_equal(ro["fm"], eo["fm"]) and np.array_equal(ro["n_active"], eo["n_active"])
        for k in KEYS:
            assert np.array_equal(rg[k], eg[k]), k
        with pytest.raises(AtacError):   # a fused graph leaves no records either
            b.vjp(g)
    # against the separate passes on this batch of uneven instances
    out = _np(b.eval(checkpoint=True))
    grads = _np(b.vjp(g))
    worst = _worst({"fm": out["fm"], **grads}, {"fm": eo["fm"], **eg})
    print(f"  vs separate passes: worst {max(worst.values()):.3g}", flush=True)
    assert max(worst.values()) <= 1e-6


def test_fused_refuses_jacobian_flags(gpu):
    w = W.peg_in_hole(n_poses=2)
    b = batch_from_workload(w)
    g = torch.zeros((b.n, 12), dtype=torch.float64, device="cuda")
    from atac import _lib as L
    from atac.batch import lib
    o = b._outputs_struct(b.alloc_outputs(True, True))
    gs = b._grad_struct(b.alloc_gradients())
    import ctypes as C
    rc = lib().atac_batch_eval_vjp(b.handle, L.EVAL_DERIV, C.byref(o), C.c_void_p(g.data_ptr()), C.byref(gs), None)
    assert rc == L.ATAC_ERR_INVALID

"""State rules of a batch's adjoint records, through the C ABI and the autograd
Function: an adjoint is refused (ATAC_ERR_STATE) when the records it would read
do not belong to the batch's current instances or have not been written yet."""
import ctypes as C

import numpy as np
import pytest
import torch

from atac import _lib as L
from atac import workloads as W
from atac.autograd import contact_wrench
from atac.batch import batch_from_workload, make_instance

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(gpu):
    w = W.peg_in_hole(n_points=1500, n_poses=6)
    return w, batch_from_workload(w)


def _insts(w):
    return [make_instance(i.grid, 0, i.points, i.E1, i.phi1, i.E2, i.phi2, i.col_th, i.kn, i.kt, i.mu, i.damping)
            for i in w.instances]


def _raw_vjp(b):
    g = torch.zeros((b.n, 12), dtype=torch.float64, device="cuda")
    grads = b.alloc_gradients()
    s = b._grad_struct(grads)
    rc = L.lib().atac_batch_vjp(b.handle, C.c_void_p(g.data_ptr()), C.byref(s),
                                C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc


def test_vjp_refused_after_set_instances(small):
    w, b = small
    print("[case] eval, set_instances, vjp", flush=True)
    b.eval()
    assert _raw_vjp(b) == L.ATAC_OK
    b.set_instances(_insts(w))
    assert _raw_vjp(b) == L.ATAC_ERR_STATE
    assert b"no evaluation" in L.lib().atac_last_error()
    b.eval()
    assert _raw_vjp(b) == L.ATAC_OK


def test_vjp_refused_after_capture_until_launch(gpu):
    """Capturing records the evaluation without running it: the adjoint has no
    records until the graph (or an eager evaluation) has run."""
    w = W.peg_in_hole(n_points=1200, n_poses=4)
    b = batch_from_workload(w)
    print("[case] capture only, then vjp", flush=True)
    out = b.alloc_outputs()
    b.graph_capture(out)
    assert _raw_vjp(b) == L.ATAC_ERR_STATE
    b.graph_launch()
    torch.cuda.synchronize()
    assert _raw_vjp(b) == L.ATAC_OK
    # the graph still evaluates the instances set after capture (it reads them at launch)
    b.set_instances(_insts(w))
    assert _raw_vjp(b) == L.ATAC_ERR_STATE
    b.graph_launch()
    torch.cuda.synchronize()
    assert _raw_vjp(b) == L.ATAC_OK


def test_vjp_after_capture_with_adjoint_in_graph(gpu):
    """A graph holding evaluation and adjoint: capture accepts the adjoint of the
    evaluation recorded before it, and replays give the eager gradients."""
    w = W.peg_in_hole(n_points=1200, n_poses=4)
    b = batch_from_workload(w)
    g = torch.tensor(np.random.default_rng(2).normal(size=(b.n, 12)), device="cuda")
    b.eval()
    eager = {k: v.clone() for k, v in b.vjp(g).items()}
    out, grads = b.alloc_outputs(), b.alloc_gradients()
    b.graph_capture(out, g_fm=g, grads=grads)
    for _ in range(3):
        for v in grads.values():
            v.zero_()
        b.graph_launch()
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(grads[k], eager[k]), k


@pytest.mark.parametrize("between", ["set_instances", "eval"])
def test_autograd_backward_raises_on_stale_records(small, between):
    w, b = small
    print("[case] autograd forward,", between, ", backward", flush=True)
    n = b.n
    E1 = torch.tensor(np.stack([i.E1 for i in w.instances]), device="cuda", requires_grad=True)
    E2 = torch.tensor(np.stack([i.E2 for i in w.instances]), device="cuda")
    phi1 = torch.tensor(np.stack([i.phi1 for i in w.instances]), device="cuda", requires_grad=True)
    phi2 = torch.tensor(np.stack([i.phi2 for i in w.instances]), device="cuda")
    kn = torch.full((n,), w.instances[0].kn, dtype=torch.float64, device="cuda", requires_grad=True)
    kt = torch.full((n,), w.instances[0].kt, dtype=torch.float64, device="cuda")
    fm = contact_wrench(b, E1, E2, phi1, phi2, kn, kt)
    if between == "set_instances":
        b.set_instances(_insts(w))
    else:
        b.eval()
    with pytest.raises(RuntimeError, match="adjoint records are gone"):
        fm.sum().backward()


def test_record_pool_bound_and_overflow(gpu):
    """The record pool bounds the active contacts of an evaluation. Within the
    bound the results equal those of a pool with a record per point; beyond it
    every output is NaN, n_active is -1 and the overflow is reported."""
    w = W.peg_in_hole(n_points=2000, n_poses=8)
    full = batch_from_workload(w, tuning={"max_active": -1})
    ref = {k: v.clone() for k, v in full.eval(p6=True).items()}
    torch.cuda.synchronize()
    used = full.workspace()
    print("[case] record pool: contacts", used, flush=True)
    assert used["capacity"] == 2000 * 8 and not used["overflowed"]
    assert used["used"] == int(ref["n_active"].sum().item()) > 100
    fit = batch_from_workload(w, tuning={"max_active": used["used"]})
    got = fit.eval(p6=True)
    torch.cuda.synchronize()
    assert fit.workspace() == {"capacity": used["used"], "used": used["used"], "overflowed": False}
    for k in ref:
        assert torch.equal(got[k], ref[k]), k
    small = batch_from_workload(w, tuning={"max_active": used["used"] // 3})
    bad = small.eval(p6=True)
    torch.cuda.synchronize()
    st = small.workspace()
    assert st["overflowed"] and st["used"] == used["used"]
    assert torch.all(bad["n_active"] == -1)
    assert torch.isnan(bad["fm"]).all() and torch.isnan(bad["dfm_dp6"]).all()
    g = small.vjp(torch.ones((small.n, 12), dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    assert torch.isnan(g["g_q1"]).all()
    with pytest.raises(L.AtacError, match="record pool"):
        small.active_ids(0)

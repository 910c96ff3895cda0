"""atac_batch_eval_instances_host: set_instances + eval_host replayed as one captured
graph. The same bits as the two eager calls, at changing poses, for each flag
combination, after the batch's buffers change (re-capture), and with the adjoint
records left for atac_batch_vjp_host."""
import ctypes as C

import numpy as np
import pytest

from atac import workloads as W
from atac._lib import lib
from atac.batch import batch_from_workload, make_instance

pytestmark = pytest.mark.gpu


def _vp(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _instances(w, shift):
    out = []
    for k, i in enumerate(w.instances):
        E1 = i.E1.copy()
        E1[:3, 3] += shift * np.array([0.3, -0.2, 0.1]) * (1 + k % 3)
        out.append(make_instance(i.grid, 0, i.points, E1, i.phi1, i.E2, i.phi2, i.col_th, i.kn, i.kt, i.mu,
                                 i.damping))
    return out


def _run(b, inst, flags, graph):
    n = b.n
    arr = (type(inst[0]) * n)(*inst)
    fm, p6, na = np.zeros(12 * n), np.zeros(12 * n), np.zeros(n, dtype=np.int32)
    deriv = bool(flags & 1)
    Km = np.zeros(144 * n) if deriv else None
    Dm = np.zeros(144 * n) if deriv else None
    p6 = p6 if flags & 3 else None
    L = lib()
    if graph:
        rc = L.atac_batch_eval_instances_host(b.handle, arr, n, flags, _vp(fm), _vp(Km), _vp(Dm), _vp(p6), _vp(na))
    else:
        rc = L.atac_batch_set_instances(b.handle, arr, n, None)
        assert rc == 0, L.atac_last_error()
        rc = L.atac_batch_eval_host(b.handle, flags, _vp(fm), _vp(Km), _vp(Dm), _vp(p6), _vp(na))
    assert rc == 0, L.atac_last_error()
    return {"fm": fm, "Km": Km, "Dm": Dm, "p6": p6, "na": na}


def test_graph_replay_matches_eager_calls(gpu):
    w = W.peg_in_hole(n_poses=10)
    b = batch_from_workload(w)
    L = lib()
    for flags in (0, 1 | 2, 8, 1 | 8):     # force; Km/Dm + p6; precise force (new buffers); precise Km/Dm
        for shift in (0.0, 0.004, -0.003):
            inst = _instances(w, shift)
            g = _run(b, inst, flags, graph=True)
            e = _run(b, inst, flags, graph=False)
            for k in g:
                if g[k] is not None:
                    assert np.array_equal(g[k], e[k]), (flags, shift, k)
            assert g["na"].sum() > 0
        # the adjoint of the replayed evaluation equals that of the eager one
        inst = _instances(w, 0.002)
        gfm = np.random.default_rng(flags).normal(size=12 * b.n)
        outs = []
        for graph in (True, False):
            _run(b, inst, flags, graph)
            gq1 = np.zeros(6 * b.n)
            rc = L.atac_batch_vjp_host(b.handle, _vp(gfm), _vp(gq1), None, None, None, None, None)
            assert rc == 0, L.atac_last_error()
            outs.append(gq1)
        assert np.array_equal(outs[0], outs[1]), flags
    # a changed instance structure is refused
    bad = _instances(w, 0.0)
    bad[0].segment = 1 - bad[0].segment if b.n > 1 else 5
    arr = (type(bad[0]) * b.n)(*bad)
    fm = np.zeros(12 * b.n)
    assert L.atac_batch_eval_instances_host(b.handle, arr, b.n, 0, _vp(fm), None, None, None, None) != 0

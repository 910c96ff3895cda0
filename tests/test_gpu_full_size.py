"""The 4,096-instance batched workload at full size (51.8M queries over 16 128^3
grids), through properties that do not depend on the size: agreement with the FP64
oracle on a sample of instances (active set outside the |d| < 1e-6 band, wrench to
1e-5 of the per-contact scale), the total active count against the oracle's
detection over every instance, bitwise determinism, graph replay equal to the eager
launch, and tuning invariance of the active set."""
import numpy as np
import pytest
import torch

from atac import workloads as W
from atac.batch import batch_from_workload
from helpers import band_mask, oracle_workload, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(out):
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().copy() for k, v in out.items()}


@pytest.fixture(scope="module")
def full(gpu):
    print("[case] batched 4096 instances, 51.8M queries", flush=True)
    w = W.batched()
    b = batch_from_workload(w)
    return w, b


def test_full_size_sample_parity(full):
    w, b = full
    out = _np(b.eval(p6=True))
    sample = list(range(0, len(w.instances), 64))       # 64 instances across the batch
    ref = oracle_workload(w, flags=O.FLAG_P6 | O.FLAG_DERIV, instances=[w.instances[i] for i in sample])
    checked = 0
    for k, i in enumerate(sample):
        r = ref[k]
        band = int(band_mask(w, w.instances[i]).sum())
        assert abs(int(out["n_active"][i]) - r["n_active"]) <= band
        if r["n_active"] and band == 0:
            assert rel_err(out["fm"][i], r["fm"], r["scale_fm"]) <= TOL
            assert rel_err(out["dfm_dp6"][i], r["dfm_dp6"], r["scale_fm"]) <= TOL
            checked += 1
    assert checked > 10


def test_full_size_active_count(full):
    """Every instance's active count against the oracle's detection (distance < 0)."""
    w, b = full
    out = _np(b.eval())
    grids = [O.GridRef.from_grid(g) for g in w.grids]
    for i, ins in enumerate(w.instances):
        pts = W.round_points_f32(w.point_sets[ins.points])
        xw = pts @ ins.E1[:3, :3].T + ins.E1[:3, 3]
        d = O.distance(grids[ins.grid], ins.E2, ins.col_th, xw)
        lo, hi = int((d < -1e-6).sum()), int((d < 1e-6).sum())
        assert lo <= int(out["n_active"][i]) <= hi, f"instance {i}"


def test_full_size_determinism_and_graph(full):
    w, b = full
    g = torch.tensor(np.random.default_rng(5).normal(size=(b.n, 12)), device="cuda")
    a1 = _np(b.eval(checkpoint=True))
    v1 = _np(b.vjp(g))
    a2 = _np(b.eval(checkpoint=True))
    v2 = _np(b.vjp(g))
    for k in a1:
        assert np.array_equal(a1[k], a2[k]), k
    for k in v1:
        assert np.array_equal(v1[k], v2[k]), k
    out = b.alloc_outputs()
    grads = b.alloc_gradients()
    b.graph_capture(out, checkpoint=True, g_fm=g, grads=grads)
    for _ in range(3):
        b.graph_launch()
        torch.cuda.synchronize()
        go, gg = _np(out), _np(grads)
        assert np.array_equal(go["fm"], a1["fm"]) and np.array_equal(go["n_active"], a1["n_active"])
        for k in v1:
            assert np.array_equal(gg[k], v1[k]), k


def test_full_size_tuning_invariance(full):
    """Tunings that keep the tile size (workgroup size, points per lane, tile order,
    persistent workgroups) leave the tiles' active lists and work items as they are,
    so every output is bitwise equal. Another tile size regroups the work items: the
    per-contact FP32 terms are the same and only their FP64 sums are reassociated,
    so the wrench agrees far inside FP32 rounding."""
    w, b = full
    base = _np(b.eval(deriv=True))
    scale = np.abs(base["fm"]).max(axis=1) + 1e-12 * np.abs(base["fm"]).max()
    worst = 0.0
    try:
        for tu in ({"detect_block": 128, "points_per_lane": 2}, {"detect_block": 1024, "points_per_lane": 1},
                   {"order": 4}, {"order": 3, "points_per_lane": 2}, {"blocks_per_cu": 2}):
            b.retune(tu)
            o = _np(b.eval(deriv=True))
            for k in base:
                assert np.array_equal(o[k], base[k]), (tu, k)
        for tu in ({"tile_points": 2048, "detect_block": 128, "points_per_lane": 2},
                   {"tile_points": 8192, "detect_block": 512, "points_per_lane": 1}):
            b.retune(tu)
            o = _np(b.eval(deriv=True))   # the same kernels as the base (FMA contraction differs between modes)
            assert np.array_equal(o["n_active"], base["n_active"])
            d = np.abs(o["fm"] - base["fm"]).max(axis=1) / scale
            worst = max(worst, float(d.max()))
            assert np.all(d <= 1e-8), tu
    finally:
        b.retune({})
    print(f"  other tile sizes: worst wrench difference {worst:.3g} of the instance's largest component", flush=True)


def test_full_size_group_culling_changes_nothing(full):
    """Dropping the groups whose boxes reach only inactive cells leaves every output
    of the evaluation and of its adjoint unchanged, bit for bit (the tiles' active
    lists are the same, in the same order); on this workload it drops a large share
    of the points."""
    w, b = full
    g = torch.tensor(np.random.default_rng(9).normal(size=(b.n, 12)), device="cuda")
    on = _np(b.eval(deriv=True))
    von = _np(b.vjp(g))
    live, total = b.live_points()
    ids_on = [b.active_ids(i) for i in range(0, b.n, 97)]
    try:
        b.retune({"group_cull": -1})
        off = _np(b.eval(deriv=True))
        voff = _np(b.vjp(g))
        live_off, _ = b.live_points()
        ids_off = [b.active_ids(i) for i in range(0, b.n, 97)]
    finally:
        b.retune({})
    print(f"  culling read {live} of {total} points ({live / total:.3f})", flush=True)
    assert live_off == total and 0 < live < 0.9 * total
    for k in on:
        assert np.array_equal(on[k], off[k]), k
    for k in von:
        assert np.array_equal(von[k], voff[k]), k
    for a, c in zip(ids_on, ids_off):
        assert np.array_equal(a, c)

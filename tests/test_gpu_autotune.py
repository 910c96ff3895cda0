"""The autotuner picks one of its candidate launch shapes by timing the step,
leaves the batch planned with it (results unchanged up to FP32 regrouping),
caches the winner per workload signature, and tunes the sort granularity."""
import json

import numpy as np
import pytest
import torch

from atac import autotune as T
from atac import workloads as W
from atac.batch import DeviceGrid, DevicePoints, ContactBatch, batch_from_workload, make_instance
from helpers import rel_err

pytestmark = pytest.mark.gpu


def test_autotune_and_cache(gpu, tmp_path):
    w = W.peg_in_hole(n_poses=32)
    b = batch_from_workload(w)
    base = b.eval()
    torch.cuda.synchronize()
    base = {k: v.cpu().numpy() for k, v in base.items()}
    cands = [{"tile_points": 1024, "detect_block": 256, "points_per_lane": 2},
             {"tile_points": 4096, "detect_block": 512, "points_per_lane": 4},
             {"tile_points": 100000}]                       # refused, skipped
    cache = str(tmp_path / "tune.json")
    tu, times = T.autotune(b, cands, reps=2, cache=cache)
    assert tu in cands[:2] and len(times) == 2
    assert b.info()["tile_points"] == tu["tile_points"]
    o = b.eval()
    torch.cuda.synchronize()
    assert np.array_equal(o["n_active"].cpu().numpy(), base["n_active"])
    scale = np.abs(base["fm"]).max()
    assert rel_err(o["fm"].cpu().numpy(), base["fm"], scale) <= 1e-6
    with open(cache) as f:
        table = json.load(f)
    assert len(table) == 1 and list(table.values())[0]["tuning"] == tu
    b.retune({"tile_points": 2048})
    tu2, _ = T.autotune(b, cands, reps=2, cache=cache)      # served from the cache
    assert tu2 == tu and b.info()["tile_points"] == tu["tile_points"]


def test_autotune_sort_granularity(gpu):
    w = W.peg_in_hole(n_poses=8)
    grids = [DeviceGrid.from_grid(g) for g in w.grids]
    pts_host = np.concatenate(w.point_sets, axis=0)
    seg = np.array([0, len(pts_host)], dtype=np.int64)

    def make(sc):
        pts = DevicePoints(pts_host, seg, sc)
        inst = [make_instance(i.grid, 0, 0, i.E1, i.phi1, i.E2, i.phi2, i.col_th, i.kn, i.kt, i.mu, i.damping)
                for i in w.instances]
        b = ContactBatch(grids, [pts], inst)
        b._keep = pts
        return b

    cands = [{"tile_points": 1024, "detect_block": 256, "points_per_lane": 2}]
    sc, tu, b = T.autotune_sort(make, [0.0, 0.05], cands, reps=2)
    assert sc in (0.0, 0.05) and tu == cands[0]
    assert int(b.eval()["n_active"].sum().item()) > 0

"""Autotuner for the batched contact path.

Picks the detection launch shape (points per tile = points per workgroup, threads
per workgroup, points per lane, tile-queue order) by timing the whole step
(evaluation + adjoint) with HIP events on the batch's own stream, and optionally the
Morton sort granularity of the point sets (which needs the point sets and the batch
rebuilt, so it takes a factory). Winners are cached in a JSON file keyed by a
signature of the workload (instances, points, grid sizes, active fraction), so a
repeated workload skips the search.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np
import torch

DEFAULT_CANDIDATES = [
    {"tile_points": tp, "detect_block": blk, "points_per_lane": ppl, "order": order}
    for order in (1, 4)
    for tp in (1024, 2048, 4096, 8192)
    for blk in (128, 256, 512)
    for ppl in (2, 4)
    if tp >= blk * ppl // 2
]


def _cache_path(path=None):
    if path:
        return path
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(base, "atac", "autotune.json")


def signature(batch, active_fraction=None):
    """Workload signature: instance count, total points, grid sizes, active fraction (2 digits)."""
    info = batch.info()
    grids = sorted({tuple(int(v) for v in g.shape) for g in batch.grids})
    sig = {"instances": batch.n, "points": info["n_points"], "grids": grids}
    if active_fraction is not None:
        sig["active"] = round(float(active_fraction), 2)
    return json.dumps(sig, sort_keys=True)


def time_step(batch, g_fm, out, grads, reps=5, warmup=2):
    """Mean milliseconds of one evaluation + adjoint."""
    for _ in range(warmup):
        batch.eval(out=out)
        batch.vjp(g_fm, out=grads)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        batch.eval(out=out)
        batch.vjp(g_fm, out=grads)
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / reps


def autotune(batch, candidates=None, reps=5, cache=None, use_cache=True, log=None):
    """Times every candidate tuning on `batch` (re-planning it each time), leaves the
    batch planned with the fastest and returns (tuning, {json(tuning): ms})."""
    candidates = candidates or DEFAULT_CANDIDATES
    g = torch.tensor(np.random.default_rng(0).normal(size=(batch.n, 12)), device="cuda")
    out = batch.alloc_outputs()
    grads = batch.alloc_gradients()
    batch.eval(out=out)
    torch.cuda.synchronize()
    n_act = int(out["n_active"].sum().item())
    key = signature(batch, n_act / max(1, batch.info()["n_points"]))
    path = _cache_path(cache)
    table = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
    if use_cache and key in table:
        best = table[key]["tuning"]
        batch.retune(best)
        return best, table[key].get("times", {})
    times = {}
    best, best_ms = None, float("inf")
    for tu in candidates:
        try:
            batch.retune(tu)
        except Exception as e:   # a shape the device refuses (LDS) is skipped
            if log:
                log(f"autotune: {tu} refused: {e}")
            continue
        ms = time_step(batch, g, out, grads, reps=reps)
        times[json.dumps(tu, sort_keys=True)] = ms
        if log:
            log(f"autotune: {tu} {ms:.4f} ms")
        if ms < best_ms:
            best, best_ms = tu, ms
    batch.retune(best)
    if use_cache:
        table[key] = {"tuning": best, "ms": best_ms, "times": times, "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    return best, times


def autotune_sort(make_batch, sort_cells, candidates=None, reps=5, log=None):
    """Sort granularity: make_batch(sort_cell) builds point sets and a batch for one
    Morton cell edge (0 = automatic); each is tuned with autotune (no cache) and the
    fastest (sort_cell, tuning, batch) is returned."""
    best = None
    for sc in sort_cells:
        b = make_batch(sc)
        tu, times = autotune(b, candidates, reps=reps, use_cache=False, log=log)
        ms = min(times.values()) if times else float("inf")
        if log:
            log(f"autotune: sort_cell {sc}: {ms:.4f} ms with {tu}")
        if best is None or ms < best[0]:
            if best is not None:
                best[3].close()
            best = (ms, sc, tu, b)
        else:
            b.close()
    return best[1], best[2], best[3]

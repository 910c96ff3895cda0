"""GPU parity of the batched contact path against the FP64 oracle.

Tolerances (relative to the FP64 reference on identical FP32-rounded inputs):
  wrench fm           ||fm - fm_ref|| <= 1e-5 * sum_c ||fm_c||   (sum over contacts of per-contact wrench norms)
  Km, Dm, p6 column   ||X - X_ref||  <= 1e-5 * sum_c ||X_c||
  active set          identical outside the |d| < 1e-6 band
The per-contact normalisation is used because peg-in-hole wrenches cancel (net
lateral force ~ 0 while each contact carries ~ kn*d); the error against ||fm_ref||
itself is reported by bench.py alongside.
"""
import numpy as np
import pytest
import torch

from atac import workloads as W
from atac.batch import ContactBatch, DeviceGrid, DevicePoints, batch_from_workload, make_instance
from helpers import band_mask, oracle_workload, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(out):
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.fixture(scope="module")
def peg(gpu):
    w = W.peg_in_hole()
    return w, oracle_workload(w)


@pytest.fixture(scope="module")
def peg_batch(peg):
    w, _ = peg
    return batch_from_workload(w)


def test_peg_in_hole_wrench_and_jacobians(peg, peg_batch):
    w, ref = peg
    print(f"[case] peg_in_hole 256 poses x 5000 points, force+Km/Dm+p6", flush=True)
    out = _np(peg_batch.eval(deriv=True))
    worst = {"fm": 0.0, "Km": 0.0, "Dm": 0.0, "p6": 0.0}
    for i, r in enumerate(ref):
        assert abs(int(out["n_active"][i]) - r["n_active"]) <= int(band_mask(w, w.instances[i]).sum())
        if r["n_active"] == 0:
            assert np.all(out["fm"][i] == 0.0) and np.all(out["Km"][i] == 0.0)
            continue
        worst["fm"] = max(worst["fm"], rel_err(out["fm"][i], r["fm"], r["scale_fm"]))
        worst["Km"] = max(worst["Km"], rel_err(out["Km"][i], r["Km"], r["scale_Km"]))
        worst["Dm"] = max(worst["Dm"], rel_err(out["Dm"][i], r["Dm"], r["scale_Dm"]))
        worst["p6"] = max(worst["p6"], rel_err(out["dfm_dp6"][i], r["dfm_dp6"], r["scale_fm"] / 1.0))
    print("worst relative errors", worst)
    for k, v in worst.items():
        assert v <= TOL, (k, v)


def test_peg_in_hole_active_set(peg, peg_batch):
    w, ref = peg
    peg_batch.eval(checkpoint=True)
    torch.cuda.synchronize()
    for i in range(0, len(ref), 16):
        ids = peg_batch.active_ids(i)
        gpu_set = np.zeros(len(w.point_sets[0]), dtype=bool)
        gpu_set[ids] = True
        diff = gpu_set != ref[i]["active"]
        band = band_mask(w, w.instances[i])
        assert not np.any(diff & ~band), f"instance {i}: {int((diff & ~band).sum())} mismatches outside the band"


def test_peg_in_hole_kt0_and_no_friction(gpu):
    for kt, mu in ((0.0, 0.3), (1e3, 0.0)):
        w = W.peg_in_hole(n_poses=16, kt=kt)
        for ins in w.instances:
            ins.mu = mu
        ref = oracle_workload(w)
        out = _np(batch_from_workload(w).eval(deriv=True))
        for i, r in enumerate(ref):
            if r["n_active"]:
                assert rel_err(out["fm"][i], r["fm"], r["scale_fm"]) <= TOL
                assert rel_err(out["Km"][i], r["Km"], r["scale_Km"]) <= TOL


def test_precise_mode(peg, peg_batch):
    w, ref = peg
    out = _np(peg_batch.eval(deriv=True, precise=True))
    for i, r in enumerate(ref):
        if r["n_active"]:
            assert rel_err(out["fm"][i], r["fm"], r["scale_fm"]) <= TOL
            assert rel_err(out["Km"][i], r["Km"], r["scale_Km"]) <= TOL


def test_determinism(peg_batch):
    a = _np(peg_batch.eval(deriv=True))
    b = _np(peg_batch.eval(deriv=True))
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_tuning_invariance(peg, peg_batch):
    import atac
    w, ref = peg
    base = _np(peg_batch.eval(p6=True))
    try:
        for tu in ({"tile_points": 256, "detect_block": 128, "points_per_lane": 1},
                   {"tile_points": 1024, "detect_block": 1024, "points_per_lane": 2},
                   {"tile_points": 8192, "detect_block": 512, "order": 1}, {"tile_points": 4096, "blocks_per_cu": 1},
                   {"order": 2, "points_per_lane": 1}, {"order": 3, "tile_points": 2048, "detect_block": 128},
                   {"order": 4}, {"order": 4, "tile_points": 1024, "points_per_lane": 2},
                   {"tile_points": 8192, "detect_block": 1024}, {"tile_points": 768, "detect_block": 64},
                   {"group_cull": -1}, {"group_cull": 1, "tile_points": 512, "detect_block": 1024},
                   {"tile_points": 2048, "blocks_per_cu": 2}):
            peg_batch.retune(tu)
            o = _np(peg_batch.eval(p6=True))
            assert np.array_equal(o["n_active"], base["n_active"])
            # the tiling regroups the FP32 per-lane sums: agreement to FP32 rounding
            for i, r in enumerate(ref):
                if r["n_active"]:
                    assert rel_err(o["fm"][i], base["fm"][i], r["scale_fm"]) <= 1e-6
        # a tuning that does not fit is refused and leaves the batch usable
        for bad in ({"tile_points": 100000}, {"tile_points": 1000}, {"points_per_lane": 3},
                    {"tile_points": 65536}, {"tile_points": 16384}, {"tile_points": 320}, {"order": 5},
                    {"group_cull": 2}):
            with pytest.raises(atac.AtacError):
                peg_batch.retune(bad)
        o = _np(peg_batch.eval(p6=True))
        assert np.array_equal(o["n_active"], base["n_active"])
    finally:
        peg_batch.retune({})


def test_plain_path_matches(peg, peg_batch):
    w, ref = peg
    a = _np(peg_batch.eval(deriv=True))
    b = _np(peg_batch.eval_plain(deriv=True))
    assert np.array_equal(a["n_active"], b["n_active"])
    for i, r in enumerate(ref):
        if r["n_active"]:
            assert rel_err(b["fm"][i], r["fm"], r["scale_fm"]) <= TOL
            assert rel_err(b["Km"][i], r["Km"], r["scale_Km"]) <= TOL


def test_graph_replay(peg, peg_batch):
    w, _ = peg
    eager = _np(peg_batch.eval(p6=True))
    out = peg_batch.alloc_outputs(p6=True)
    peg_batch.graph_capture(out, p6=True)
    for rep in range(3):
        for v in out.values():
            v.zero_()
        peg_batch.graph_launch()
        got = _np(out)
        for k in eager:
            assert np.array_equal(got[k], eager[k]), (rep, k)
        torch.ones(1000, device="cuda").sum().item()  # eager work between replays


def test_set_instances_updates_poses(peg, peg_batch):
    w, ref = peg
    rot = list(reversed(w.instances))
    inst = [make_instance(i.grid, 0, i.points, i.E1, i.phi1, i.E2, i.phi2, i.col_th, i.kn, i.kt, i.mu, i.damping)
            for i in rot]
    peg_batch.set_instances(inst)
    out = _np(peg_batch.eval())
    n = len(ref)
    for i in range(0, n, 8):
        r = ref[n - 1 - i]
        if r["n_active"]:
            assert rel_err(out["fm"][i], r["fm"], r["scale_fm"]) <= TOL
    inst = [make_instance(i.grid, 0, i.points, i.E1, i.phi1, i.E2, i.phi2, i.col_th, i.kn, i.kt, i.mu, i.damping)
            for i in w.instances]
    peg_batch.set_instances(inst)


def test_blobs_parity(gpu):
    print("[case] blobs 2 pairs x 50k points x 8 poses", flush=True)
    w = W.blobs(n_pairs=2, n_points=50_000, n_poses=8)
    ref = oracle_workload(w)
    out = _np(batch_from_workload(w).eval(deriv=True))
    assert sum(r["n_active"] for r in ref) > 0
    for i, r in enumerate(ref):
        band = int(band_mask(w, w.instances[i]).sum())
        assert abs(int(out["n_active"][i]) - r["n_active"]) <= band
        if r["n_active"] and band == 0:
            assert rel_err(out["fm"][i], r["fm"], r["scale_fm"]) <= TOL
            assert rel_err(out["Km"][i], r["Km"], r["scale_Km"]) <= TOL
            assert rel_err(out["Dm"][i], r["Dm"], r["scale_Dm"]) <= TOL


def test_batched_mixed_parts(gpu):
    print("[case] batched 48 instances, mixed parts, 1k-20k points", flush=True)
    w = W.batched(n_inst=48, n_max=20_000, grid_n=64)
    ref = oracle_workload(w, flags=O.FLAG_P6 | O.FLAG_DERIV)
    out = _np(batch_from_workload(w, tuning={"tile_points": 2048}).eval(p6=True))
    for i, r in enumerate(ref):
        band = int(band_mask(w, w.instances[i]).sum())
        assert abs(int(out["n_active"][i]) - r["n_active"]) <= band
        if r["n_active"] and band == 0:
            assert rel_err(out["fm"][i], r["fm"], r["scale_fm"]) <= TOL
            assert rel_err(out["dfm_dp6"][i], r["dfm_dp6"], r["scale_fm"]) <= TOL


def test_edge_cases(gpu):
    """Empty point set, all points outside the box, a single point, shared grids."""
    w = W.peg_in_hole(n_points=64, n_poses=2)
    g = DeviceGrid.from_grid(w.grids[0])
    pts = W.round_points_f32(w.point_sets[0])
    far = pts + 10.0
    allp = np.concatenate([pts, far, pts[:1]])
    seg = np.array([0, 64, 64, 128, 129], dtype=np.int64)   # segments: pts, empty, far, one point
    P = DevicePoints(allp, seg)
    I = w.instances[0]
    E1 = np.eye(4)
    E1[:3, 3] = [0.04, 0.0, 0.0]
    inst = [make_instance(0, 0, s, E1, I.phi1, I.E2, I.phi2, I.col_th, I.kn, I.kt, I.mu, I.damping) for s in range(4)]
    b = ContactBatch([g], [P], inst)
    out = _np(b.eval(deriv=True))
    G = O.GridRef.from_grid(w.grids[0])
    for s in range(4):
        sub = allp[seg[s]:seg[s + 1]]
        r = O.eval_instance([G], O.make_instance(0, sub, E1, I.phi1, I.E2, I.phi2, I.col_th, I.kn, I.kt, I.mu,
                                                 I.damping), O.FLAG_DERIV)
        assert int(out["n_active"][s]) == r["n_active"]
        if r["n_active"]:
            assert rel_err(out["fm"][s], r["fm"], r["scale_fm"]) <= TOL
        else:
            assert np.all(out["fm"][s] == 0.0)
    assert out["n_active"][1] == 0 and out["n_active"][2] == 0


def test_touched_values(gpu):
    w = W.peg_in_hole(n_points=3000, n_poses=4)
    b = batch_from_workload(w)
    got = b.touched_values()
    g = w.grids[0]
    ni, nj, nk = g.n
    seen = set()
    pts = W.round_points_f32(w.point_sets[0]).astype(np.float32).astype(np.float64)
    for ins in w.instances:
        A = ins.E2[:3, :3].T @ ins.E1[:3, :3]
        t = ins.E2[:3, :3].T @ (ins.E1[:3, 3] - ins.E2[:3, 3])
        c = (pts @ A.T + t - g.min_box) / g.dxyz
        inside = np.all((c >= 0) & (c <= (g.max_box - g.min_box) / g.dxyz), axis=1)
        i0 = np.minimum(np.floor(c[inside]).astype(np.int64), np.array([ni, nj, nk]) - 1)
        i1 = np.minimum(i0 + 1, np.array([ni, nj, nk]) - 1)
        for cx in (0, 1):
            for cy in (0, 1):
                for cz in (0, 1):
                    ii = np.where(cx, i1[:, 0], i0[:, 0])
                    jj = np.where(cy, i1[:, 1], i0[:, 1])
                    kk = np.where(cz, i1[:, 2], i0[:, 2])
                    seen.update((ii + ni * (jj + nj * kk)).tolist())
    assert abs(got - len(seen)) <= 0.001 * len(seen)


def test_distance_api(gpu):
    w = W.peg_in_hole(n_points=4000, n_poses=1)
    g = w.grids[0]
    rng = np.random.default_rng(5)
    xw = rng.uniform(-1.7, 1.7, size=(4000, 3))
    E2 = W.pose(W.rotation_exp([0.1, -0.2, 0.3]), [0.05, 0.0, -0.02])
    got = DeviceGrid.from_grid(g).distance(xw, E2, 0.01).cpu().numpy()
    ref = O.distance(O.GridRef.from_grid(g), E2, 0.01, xw)
    inside = ref < 1e30
    assert np.array_equal(got >= 1e30, ~inside)
    assert np.max(np.abs(got[inside] - ref[inside])) <= 1e-6


# BodySDFObj::collision values and derivative blocks against the FP64 oracle, relative to
# the largest entry: values to 1e-5 (FP64 arithmetic) and 2e-6 (FP32: about 16 units of
# FP32 rounding; 1.4e-7 measured on the host), blocks to 1e-4 in both (the closed-form
# cell derivatives against the oracle's FD stencils: 7.9e-7 measured)
@pytest.mark.parametrize("precise,tol_v,tol_d", [(True, 1e-5, 1e-4), (False, 2e-6, 1e-4)])
def test_collision_api(gpu, precise, tol_v, tol_d):
    w = W.peg_in_hole(n_points=2000, n_poses=1)
    g = w.grids[0]
    rng = np.random.default_rng(6)
    xw = rng.uniform(-1.2, 1.2, size=(500, 3))
    xd = rng.normal(size=(500, 3))
    E2 = W.pose(W.rotation_exp([0.05, 0.02, -0.1]), [0.01, 0.02, 0.0])
    phi2 = rng.normal(0, 0.3, 6)
    G = O.GridRef.from_grid(g)
    ref = O.collision(G, E2, phi2, 0.01, xw, xd, derivs=True)
    keep = np.abs(ref["d"]) < 1.0          # inside the box: gradient defined
    vals, der = DeviceGrid.from_grid(g).collision(xw, xd, E2, phi2, 0.01, derivs=True, precise=precise)
    vals, der = vals.cpu().numpy(), der.cpu().numpy()
    refv = np.concatenate([ref["d"][:, None], ref["n"], ref["ddot"][:, None], ref["tdot"], ref["xi2"]], axis=1)
    ev = np.max(np.abs(vals[keep] - refv[keep])) / max(1.0, np.abs(refv[keep]).max())
    off, worst = 0, {}
    for name, shape in O.COLLISION_DERIV_LAYOUT:
        cnt = int(np.prod(shape))
        a = der[keep, off:off + cnt]
        bref = ref[name][keep].reshape(-1, cnt)
        worst[name] = np.max(np.abs(a - bref)) / max(1.0, np.abs(bref).max())
        off += cnt
    print(f"  precise={precise}: {int(keep.sum())} points, values {ev:.3g}, blocks worst "
          f"{max(worst.values()):.3g} ({max(worst, key=worst.get)})", flush=True)
    assert ev <= tol_v
    for name, e in worst.items():
        assert e <= tol_d, name


def test_nonfinite_nodes(gpu):
    """Grid nodes holding NaN or +inf (outside the part, where the field is
    positive) leave their blocks open: the points there are interpolated, stay
    inactive as in the oracle, and the wrench of the active points is unchanged."""
    import dataclasses
    w = W.peg_in_hole(n_points=4000, n_poses=8)
    g0 = w.grids[0]
    vals = np.array(g0.values, dtype=np.float32, copy=True)
    flat = vals.reshape(-1)
    rng = np.random.default_rng(21)
    far = np.flatnonzero(flat > 0.1)
    pick = rng.choice(far, size=600, replace=False)
    flat[pick[:300]] = np.nan
    flat[pick[300:]] = np.inf
    wp = dataclasses.replace(w, grids=[dataclasses.replace(g0, values=vals)])
    print(f"[case] peg_in_hole 8 poses x 4000 points, {len(pick)} non-finite nodes", flush=True)
    ref = oracle_workload(wp, flags=O.FLAG_DERIV)
    b = batch_from_workload(wp)
    out = _np(b.eval(deriv=True, checkpoint=True))
    n_checked = 0
    for i, r in enumerate(ref):
        band = band_mask(wp, wp.instances[i])
        ids = b.active_ids(i)
        gpu_set = np.zeros(len(wp.point_sets[0]), dtype=bool)
        gpu_set[ids] = True
        assert not np.any((gpu_set != r["active"]) & ~band), i
        if r["n_active"] and not band.any():
            assert np.all(np.isfinite(out["fm"][i]))
            assert rel_err(out["fm"][i], r["fm"], r["scale_fm"]) <= TOL
            n_checked += 1
    assert n_checked > 0


def test_ragged_segment_sizes(gpu):
    """Point sets of sizes on both sides of the wavefront (64), step (256 with four
    points per lane) and tile (2048) boundaries, each the prefix of one contact
    point set, in one batch: active counts and wrench equal the oracle's per set."""
    w = W.peg_in_hole(n_points=5000, n_poses=2)
    pts = W.round_points_f32(w.point_sets[0])
    sizes = [1, 63, 64, 65, 255, 256, 257, 2047, 2048, 2049, 4095, 4097]
    allp = np.concatenate([pts[:n] for n in sizes])
    seg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    P = DevicePoints(allp, seg)
    g = DeviceGrid.from_grid(w.grids[0])
    I = w.instances[1]
    print(f"[case] ragged segments {sizes}", flush=True)
    inst = [make_instance(0, 0, s, I.E1, I.phi1, I.E2, I.phi2, I.col_th, I.kn, I.kt, I.mu, I.damping)
            for s in range(len(sizes))]
    b = ContactBatch([g], [P], inst, tuning={"tile_points": 2048, "points_per_lane": 4})
    out = _np(b.eval(deriv=True))
    G = O.GridRef.from_grid(w.grids[0])
    n_contact = 0
    for s, n in enumerate(sizes):
        r = O.eval_instance([G], O.make_instance(0, pts[:n], I.E1, I.phi1, I.E2, I.phi2, I.col_th, I.kn, I.kt,
                                                 I.mu, I.damping), O.FLAG_DERIV)
        assert int(out["n_active"][s]) == r["n_active"], (n, int(out["n_active"][s]), r["n_active"])
        if r["n_active"]:
            n_contact += 1
            assert rel_err(out["fm"][s], r["fm"], r["scale_fm"]) <= TOL, n
            assert rel_err(out["Km"][s], r["Km"], r["scale_Km"]) <= TOL, n
        else:
            assert np.all(out["fm"][s] == 0.0)
    assert n_contact >= 6

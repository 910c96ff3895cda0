"""CPU checks of the contact library's arithmetic (contact_math.h, compiled for the
host) against the FP64 oracle: detection, wrench, Km/Dm tangent columns, the p6
column, the hand-written adjoint (against the oracle's Km/Dm and against central
finite differences), and the FP64 slow path near cell faces and the box boundary."""
import numpy as np
import pytest

from atac import workloads as W
from atac.batch import make_instance
from helpers import band_mask, oracle_workload, rel_err
from hostcheck import hostcheck as H
from oracle import oracle as O

TOL = 1e-5


def _ai(ins):
    return make_instance(ins.grid, 0, 0, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt, ins.mu,
                         ins.damping, ins.scale, ins.contact_scale)


@pytest.fixture(scope="module")
def peg():
    w = W.peg_in_hole(n_poses=48)
    return w, oracle_workload(w)


@pytest.mark.parametrize("precise", [False, True])
def test_peg_wrench_jacobians_p6(peg, precise):
    w, ref = peg
    pts = W.round_points_f32(w.point_sets[0])
    for i, r in enumerate(ref):
        h = H.eval_instance(w.grids[0], pts, _ai(w.instances[i]), deriv=True, precise=precise)
        diff = h["active"] != r["active"]
        assert not np.any(diff & ~band_mask(w, w.instances[i]))
        if not r["n_active"]:
            assert np.all(h["fm"] == 0)
            continue
        assert rel_err(h["fm"], r["fm"], r["scale_fm"]) <= TOL
        assert rel_err(h["Km"], r["Km"], r["scale_Km"]) <= TOL
        assert rel_err(h["Dm"], r["Dm"], r["scale_Dm"]) <= TOL
        assert rel_err(h["dfm_dp6"], r["dfm_dp6"], r["scale_fm"]) <= TOL


def test_vjp_equals_jacobian_contraction(peg):
    w, ref = peg
    pts = W.round_points_f32(w.point_sets[0])
    rng = np.random.default_rng(4)
    for i, r in enumerate(ref[:16]):
        if not r["n_active"]:
            continue
        g = rng.normal(size=12)
        h = H.eval_instance(w.grids[0], pts, _ai(w.instances[i]), g_fm=g)
        s = np.linalg.norm(g) * r["scale_Km"]
        assert rel_err(h["vjp"][:12], g @ r["Km"], s) <= TOL
        assert rel_err(h["vjp"][12:24], g @ r["Dm"], np.linalg.norm(g) * r["scale_Dm"]) <= TOL


def _fixed_fm(w, i, ids, E1=None, E2=None, kn=None, kt=None, phi1=None, phi2=None):
    ins = w.instances[i]
    G = O.GridRef.from_grid(w.grids[ins.grid])
    pts = W.round_points_f32(w.point_sets[ins.points])[ids]
    return O.force_fixed_contacts(G, pts, ins.E1 if E1 is None else E1, ins.phi1 if phi1 is None else phi1,
                                  ins.E2 if E2 is None else E2, ins.phi2 if phi2 is None else phi2, ins.col_th,
                                  ins.kn if kn is None else kn, ins.kt if kt is None else kt, ins.mu,
                                  ins.damping)["fm"]


def _multilinear(mu, kt):
    import test_gpu_adjoint
    return test_gpu_adjoint.multilinear_workload(mu=mu, kt=kt)


@pytest.mark.parametrize("mu,kt", [(0.3, 1e3), (0.0, 1e3)], ids=["friction", "frictionless"])
def test_vjp_central_fd_h1e4_multilinear(mu, kt):
    """The hand-written adjoint (host-compiled) against FP64 central FD with h = 1e-4 on a
    multilinear field, which trilinear interpolation reproduces in every cell, over the
    contacts whose friction branch cannot switch inside the step."""
    from helpers import branch_stable_contacts
    w = _multilinear(mu, kt)
    ref = oracle_workload(w)
    pts = W.round_points_f32(w.point_sets[0])
    rng = np.random.default_rng(9)
    h = 1e-4
    n_static = 0
    for i, r in enumerate(ref[:4]):
        ins = w.instances[i]
        ids = branch_stable_contacts(w, ins, np.nonzero(r["active"])[0])
        g = rng.normal(size=12)
        sub = O.force_fixed_contacts(O.GridRef.from_grid(w.grids[0]), pts[ids], ins.E1, ins.phi1, ins.E2, ins.phi2,
                                     ins.col_th, ins.kn, ins.kt, ins.mu, ins.damping, derivs=True)
        hv = H.eval_instance(w.grids[0], pts[ids], _ai(ins), g_fm=g, precise=True)
        assert hv["n_active"] == len(ids)
        v = hv["vjp"]
        fd = np.zeros(24)
        for k in range(6):
            dq = np.zeros(6)
            dq[k] = h
            fd[k] = g @ (_fixed_fm(w, i, ids, E1=ins.E1 @ O.se3_exp(dq)) - _fixed_fm(w, i, ids, E1=ins.E1 @ O.se3_exp(-dq))) / (2 * h)
            fd[6 + k] = g @ (_fixed_fm(w, i, ids, E2=ins.E2 @ O.se3_exp(dq)) - _fixed_fm(w, i, ids, E2=ins.E2 @ O.se3_exp(-dq))) / (2 * h)
            fd[12 + k] = g @ (_fixed_fm(w, i, ids, phi1=ins.phi1 + dq) - _fixed_fm(w, i, ids, phi1=ins.phi1 - dq)) / (2 * h)
            fd[18 + k] = g @ (_fixed_fm(w, i, ids, phi2=ins.phi2 + dq) - _fixed_fm(w, i, ids, phi2=ins.phi2 - dq)) / (2 * h)
        sK = np.linalg.norm(g) * np.abs(sub["Km"]).sum()
        sD = np.linalg.norm(g) * max(np.abs(sub["Dm"]).sum(), 1e-30)
        assert rel_err(v[:12], fd[:12], sK) <= TOL
        assert rel_err(v[12:24], fd[12:24], sD) <= TOL
        assert rel_err(v[:12], g @ sub["Km"], sK) <= TOL
        sf = np.linalg.norm(g) * np.abs(sub["fm"]).sum()
        dk = ins.kn * 1e-5
        fdk = g @ (_fixed_fm(w, i, ids, kn=ins.kn + dk) - _fixed_fm(w, i, ids, kn=ins.kn - dk)) / (2 * dk)
        assert abs(v[24] - fdk) <= 1e-7 * sf / ins.kn
        dt = ins.kt * 1e-5
        fdt = g @ (_fixed_fm(w, i, ids, kt=ins.kt + dt) - _fixed_fm(w, i, ids, kt=ins.kt - dt)) / (2 * dt)
        assert abs(v[25] - fdt) <= 1e-7 * sf / ins.kt
        n_static += len(ids)
    assert n_static > 100


@pytest.mark.parametrize("precise,tol_v", [(True, 1e-5), (False, 2e-6)])
def test_collision_math_matches_oracle(precise, tol_v):
    """collision_kernel's arithmetic (BodySDFObj::collision, BodySDFObj.cpp:251-376) on the
    host, through the instance record atac_sdf_collision builds (body 1 the world frame),
    against the FP64 oracle: values relative to their largest entry to 1e-5 (FP64) and
    2e-6 (FP32), derivative blocks to 1e-4 (the closed-form cell derivatives against the
    oracle's FD stencils). A record without its FP32 copies of the poses (R2 = 0 for the
    FP32 arithmetic) gives errors of order 1."""
    w = W.peg_in_hole(n_points=2000, n_poses=1)
    g = w.grids[0]
    rng = np.random.default_rng(6)
    xw = rng.uniform(-1.2, 1.2, size=(500, 3))
    xd = rng.normal(size=(500, 3))
    E2 = W.pose(W.rotation_exp([0.05, 0.02, -0.1]), [0.01, 0.02, 0.0])
    phi2 = rng.normal(0, 0.3, 6)
    ref = O.collision(O.GridRef.from_grid(g), E2, phi2, 0.01, xw, xd, derivs=True)
    keep = np.abs(ref["d"]) < 1.0
    refv = np.concatenate([ref["d"][:, None], ref["n"], ref["ddot"][:, None], ref["tdot"], ref["xi2"]], axis=1)
    vals, der = H.collision(g, E2, phi2, 0.01, xw, xd, precise=precise)
    assert keep.sum() > 100
    assert np.max(np.abs(vals[keep] - refv[keep])) <= tol_v * max(1.0, np.abs(refv[keep]).max())
    off = 0
    for name, shape in O.COLLISION_DERIV_LAYOUT:
        cnt = int(np.prod(shape))
        bref = ref[name][keep].reshape(-1, cnt)
        assert np.max(np.abs(der[keep, off:off + cnt] - bref)) <= 1e-4 * max(1.0, np.abs(bref).max()), name
        off += cnt


def test_slow_path_near_faces_and_box():
    """Points placed within the FD step of cell faces, cell corners and the box boundary
    take the FP64 stencil path and reproduce the reference's FD blend exactly."""
    w = W.peg_in_hole(n_points=16, n_poses=1)
    g = w.grids[0]
    rng = np.random.default_rng(8)
    base = -1.575 + 0.05 * rng.integers(4, 60, size=(3000, 3))
    off = rng.choice([-3e-6, -1e-6, 0.0, 1e-6, 3e-6, 7e-6, 0.0125, 0.025], size=(3000, 3))
    x = (base + off).astype(np.float32).astype(np.float64)
    G = O.GridRef.from_grid(g)
    E1 = np.eye(4)
    ins = W.Instance(0, 0, E1, np.zeros(6), np.eye(4), np.zeros(6), col_th=-5.0)   # d = phi - 5: every inside point active
    ref = O.eval_instance([G], O.make_instance(0, x, E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt,
                                               ins.mu, ins.damping), O.FLAG_DERIV)
    h = H.eval_instance(g, x.astype(np.float32), _ai(ins), deriv=True)
    assert h["slow"].sum() > 500
    assert h["n_active"] == ref["n_active"]
    assert rel_err(h["fm"], ref["fm"], ref["scale_fm"]) <= TOL
    assert rel_err(h["Km"], ref["Km"], ref["scale_Km"]) <= TOL


def _decisions_agree(grid, pts, inst):
    dec, full, settled, batched, slow, slow_ref = H.detect_decisions(grid, pts, inst)
    assert np.array_equal(dec, full), f"{int((dec != full).sum())} bounded decisions differ from the full ones"
    assert np.array_equal(batched, full), f"{int((batched != full).sum())} batched decisions differ"
    assert np.array_equal(slow, slow_ref), f"{int((slow != slow_ref).sum())} slow-path flags differ"
    assert not np.any(slow & ~full)
    return settled


def test_bounded_detection_matches_full_interpolation(peg):
    """The block-bound shortcut of the detection test decides every point exactly
    as the full 8-corner FP32 interpolation (on real poses, including points
    deep inside, outside the box and near the surface)."""
    w, _ = peg
    pts = W.round_points_f32(w.point_sets[0]).astype(np.float32)
    settled = []
    for ins in w.instances[:16]:
        settled.append(_decisions_agree(w.grids[0], pts, _ai(ins)).mean())
    assert 0.0 < np.mean(settled) < 1.0


def test_bounded_detection_adversarial():
    """Fields whose value sits at, or within a few FP32 ulps of, the threshold -col_th
    (where the bound must defer to the interpolation), non-finite corners, and
    points on the box faces."""
    rng = np.random.default_rng(5)
    col_th = 0.01
    base = np.float32(-col_th)
    n = 9
    zs = np.zeros(6)
    ins = make_instance(0, 0, 0, np.eye(4), zs, np.eye(4), zs, col_th, 1e4, 0.0, 0.0, 0.0)
    # points: uniform in the box, on every face and corner, and on the max face exactly
    lo, hi = -1.0, 1.0
    p = rng.uniform(lo, hi, (4000, 3))
    faces = rng.uniform(lo, hi, (600, 3))
    for a in range(3):
        faces[a * 200:(a * 200 + 100), a] = lo
        faces[(a * 200 + 100):(a + 1) * 200, a] = hi
    pts = np.concatenate([p, faces, np.array([[hi, hi, hi], [lo, lo, lo], [hi, lo, hi]])]).astype(np.float32)
    dxyz = (hi - lo) / (n - 1)
    cases = []
    for k in (-3, -1, 0, 1, 3):       # constant fields a few ulps either side of the threshold
        cases.append(np.full((n, n, n), base + np.float32(k) * np.spacing(base), dtype=np.float32))
    noisy = (base + rng.normal(0, 1e-7, (n, n, n))).astype(np.float32)
    cases.append(noisy)
    mixed = (base + rng.normal(0, 0.02, (n, n, n))).astype(np.float32)
    cases.append(mixed)
    bad = mixed.copy()
    bad[2, 3, 4] = np.nan
    bad[5, 5, 5] = np.inf
    bad[0, 0, 0] = -np.inf
    cases.append(bad)
    for vals in cases:
        g = W.SDFGrid(vals, np.array([lo] * 3), np.array([hi] * 3), np.array([dxyz] * 3))
        _decisions_agree(g, pts, ins)


def test_bounded_detection_box_inside_and_beyond_grid():
    """max_box below the grid's last node (the box face cuts interior cells, whose
    table entries defer to the exact test) and beyond it (points past the last node
    are inside and take the clamped last cell): the table's decisions equal the
    full interpolation's, for points in and around the box."""
    rng = np.random.default_rng(9)
    n, lo = 11, -1.0
    dx = 0.2
    col_th = 0.01
    zs = np.zeros(6)
    ins = make_instance(0, 0, 0, np.eye(4), zs, np.eye(4), zs, col_th, 1e4, 0.0, 0.0, 0.0)
    xs = lo + dx * np.arange(n)
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    field = (0.35 * X + 0.2 * Y - 0.3 * Z + 0.05 + 0.02 * rng.normal(size=X.shape)).transpose(2, 1, 0)
    vals = np.ascontiguousarray(field, dtype=np.float32)
    for hi in (lo + 6.3 * dx, lo + (n - 1) * dx, lo + (n - 1) * dx + 0.37):
        p = rng.uniform(lo - 0.3, hi + 0.3, (6000, 3))
        on_face = rng.uniform(lo, hi, (600, 3))
        on_face[:300, 0] = hi
        on_face[300:, 2] = lo
        cell_faces = lo + dx * rng.integers(0, n, (400, 3)) + rng.uniform(0, 1, (400, 3)) * np.array([0, 1, 1]) * dx
        pts = np.concatenate([p, on_face, cell_faces]).astype(np.float32)
        g = W.SDFGrid(vals, np.array([lo] * 3), np.array([hi] * 3), np.array([dx] * 3))
        settled = _decisions_agree(g, pts, ins)
        assert 0.0 < settled.mean() < 1.0


def _morton_order(p):
    lo = p.min(0)
    ext = max(float((p.max(0) - lo).max()), 1e-30)
    q = np.minimum(((p - lo) / ext * 1024).astype(np.int64), 1023)
    key = np.zeros(len(p), dtype=np.int64)
    for b in range(10):
        for a in range(3):
            key |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return np.argsort(key, kind="stable")


def _cull_checks(grid, pts, inst, gsize=256):
    live, needed, exact = H.group_cull(grid, pts, inst, gsize)
    assert exact.all(), f"{int((~exact).sum())} summed-area counts differ from the block-by-block count"
    assert not np.any(needed & ~live), f"{int((needed & ~live).sum())} groups dropped with an active point"
    return live, needed


def test_group_culling_keeps_every_group_that_can_hold_a_candidate(peg):
    """Group culling on real poses (Morton-sorted groups of 256, as the library cuts
    them, and of 16): a group is dropped only when none of its points is active
    (by the full interpolation), and the summed-area table's counts are the
    blocks' own."""
    w, _ = peg
    pts = W.round_points_f32(w.point_sets[0]).astype(np.float32)
    pts = pts[_morton_order(pts)]
    dropped = []
    for ins in w.instances[:16]:
        _cull_checks(w.grids[0], pts, _ai(ins), 256)
        # the peg lies along the hole's surface: groups of 16 points are needed to drop some
        live, needed = _cull_checks(w.grids[0], pts, _ai(ins), 16)
        dropped.append(1.0 - live.mean())
    print(f"  dropped groups of 16: mean {np.mean(dropped):.3f}")
    assert 0.0 < np.mean(dropped) < 1.0


def test_group_culling_adversarial():
    """Fields at or within a few ulps of the threshold, non-finite nodes, points on
    cell faces and box faces, outside the box, and groups holding a non-finite point
    (never dropped); groups of 1, 7 and 256 points."""
    rng = np.random.default_rng(7)
    col_th = 0.01
    base = np.float32(-col_th)
    n = 9
    lo, hi = -1.0, 1.0
    dxyz = (hi - lo) / (n - 1)
    zs = np.zeros(6)
    # points on the cell faces (multiples of the spacing), inside, outside the box
    faces = (np.round(rng.uniform(lo, hi, (900, 3)) / dxyz) * dxyz)
    inside = rng.uniform(lo, hi, (1200, 3))
    outside = rng.uniform(2 * lo, 2 * hi, (600, 3))
    pts = np.concatenate([faces, inside, outside]).astype(np.float32)
    pts = pts[_morton_order(pts.astype(np.float64))]
    bad_pts = pts.copy()
    bad_pts[5] = np.nan
    bad_pts[700, 1] = np.inf
    cases = [np.full((n, n, n), base + np.float32(k) * np.spacing(base), dtype=np.float32) for k in (-3, 0, 3)]
    mixed = (base + rng.normal(0, 0.3, (n, n, n))).astype(np.float32)
    cases.append(mixed)
    far = np.full((n, n, n), 1.0, dtype=np.float32)
    far[4, 4, 4] = -1.0
    cases.append(far)
    bad = mixed.copy()
    bad[2, 3, 4] = np.nan
    bad[5, 5, 5] = np.inf
    cases.append(bad)
    for vals in cases:
        g = W.SDFGrid(vals, np.array([lo] * 3), np.array([hi] * 3), np.array([dxyz] * 3))
        for E in (np.eye(4), W.pose(W.rotation_exp(np.array([0.3, -0.2, 0.5])), np.array([0.05, 0.1, -0.07]))):
            ins = make_instance(0, 0, 0, E, zs, np.eye(4), zs, col_th, 1e4, 0.0, 0.0, 0.0)
            for gs in (1, 7, 256):
                _cull_checks(g, pts, ins, gs)
            live, _ = _cull_checks(g, bad_pts, ins, 256)
            assert live[5 // 256] and live[700 // 256]


def test_group_culling_box_inside_and_beyond_grid():
    """Group culling where max_box is below the grid's last node, on it and beyond it
    (only on it is the layer beyond the last node judged by its interior neighbour),
    with points in and around the box and on its max faces: no group holding an
    active point is dropped; groups of 1, 16 and 256 points."""
    rng = np.random.default_rng(11)
    n, lo = 11, -1.0
    dx = 0.2
    col_th = 0.01
    zs = np.zeros(6)
    ins = make_instance(0, 0, 0, np.eye(4), zs, np.eye(4), zs, col_th, 1e4, 0.0, 0.0, 0.0)
    xs = lo + dx * np.arange(n)
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    field = (0.35 * X + 0.2 * Y - 0.3 * Z + 0.05 + 0.02 * rng.normal(size=X.shape)).transpose(2, 1, 0)
    vals = np.ascontiguousarray(field, dtype=np.float32)
    for hi in (lo + 6.3 * dx, lo + (n - 1) * dx, lo + (n - 1) * dx + 0.37):
        p = rng.uniform(lo - 0.3, hi + 0.3, (3000, 3))
        on_face = rng.uniform(lo, hi, (900, 3))
        for a in range(3):
            on_face[300 * a:300 * (a + 1), a] = hi
        pts = np.concatenate([p, on_face]).astype(np.float32)
        pts = pts[_morton_order(pts.astype(np.float64))]
        g = W.SDFGrid(vals, np.array([lo] * 3), np.array([hi] * 3), np.array([dx] * 3))
        for gs in (1, 16, 256):
            live, needed = _cull_checks(g, pts, ins, gs)
            assert needed.any()

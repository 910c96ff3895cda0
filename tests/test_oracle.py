"""Pins the FP64 oracle against what the mathematics fixes (the reference ships no
golden vectors or fixtures for this path, and its sources cannot be linked here):

  query_SDF          exact on multilinear fields (trilinear reproduces them), the
                     i == ni-1 clamp, FLT_MAX outside, the outside_accurate rule
  query_dSDF         = closed-form partials inside a cell; the FD blend of two cells'
                     slopes within eps of a face (closed form)
  query_ddSDF        diagonal ~ 0 inside a cell, the (sR - sL)(h - delta)/h^2 spike at a
                     face, mixed terms = bilinear cross coefficients
  math::exp          = scipy.linalg.expm of the 4x4 twist matrix
  collision          derivative blocks = central FD of the values (analytic mode)
  force              hand-computed single contact on a plane; world force balance;
                     world torque balance of the normal pair; kn linearity; kt = 0
  Km / Dm            = central FD of fm with the contact set held fixed
  p6 column          = FD in contact_scale when damping = 0; the reference's formula is
                     kept (and differs from FD) for dynamic friction with damping
Each check is chosen so a dropped term, a wrong sign or a transposed operand fails it.
"""
import numpy as np
import pytest
import scipy.linalg

from atac import workloads as W
from oracle import oracle as O

FLT_MAX = float(np.finfo(np.float32).max)


def grid_from(fn, n=(12, 11, 10), lo=(-0.5, -0.4, -0.3), d=(0.1, 0.09, 0.08)):
    g = W.make_grid(fn, n, lo, d)
    return g, O.GridRef.from_grid(g)


def rand_inside(g, rng, m, margin=0.0):
    lo = g.min_box + margin
    hi = g.max_box - margin
    return rng.uniform(lo, hi, size=(m, 3))


# ---------------------------------------------------------------- query_SDF
def test_trilinear_reproduces_multilinear_fields():
    rng = np.random.default_rng(0)
    fields = [lambda x, y, z: 0.3 * x - 0.2 * y + 0.7 * z + 0.1,
              lambda x, y, z: 0.5 * x * y * z - 0.25 * x * y + 0.125 * z - 0.05]
    for f in fields:
        g, G = grid_from(lambda x, y, z: f(x, y, z).astype(np.float32).astype(np.float64))
        x = rand_inside(g, rng, 500)
        got = O.query_sdf(G, x)
        # grid values are FP32-rounded; the interpolant of the rounded nodes is within 1e-7
        assert np.max(np.abs(got - f(x[:, 0], x[:, 1], x[:, 2]))) < 2e-7


def test_outside_returns_flt_max_and_accurate_rule():
    g, G = grid_from(lambda x, y, z: 0.3 * x + 0.2 * y - 0.1 * z)
    x = np.array([[10.0, 0.0, 0.0], [0.0, -5.0, 0.1], [g.max_box[0] + 0.3, g.max_box[1] + 0.4, 0.0]])
    assert np.all(O.query_sdf(G, x) == FLT_MAX)
    acc = O.query_sdf(G, x, outside_accurate=True)
    xb = np.clip(x, g.min_box, g.max_box)
    expect = np.linalg.norm(x - xb, axis=1) + O.query_sdf(G, xb)
    assert np.allclose(acc, expect, rtol=0, atol=1e-12)
    assert abs(acc[2] - (0.5 + O.query_sdf(G, xb[2:3])[0])) < 1e-12   # 0.3-0.4-0.5 triangle


def test_last_node_clamp():
    g, G = grid_from(lambda x, y, z: x * 0 + y * 0 + z * 0 + 1.0)
    vals = g.values.copy()
    vals[-1, -1, -1] = 7.0
    G2 = O.GridRef(vals, g.min_box, g.max_box, g.dxyz)
    assert O.query_sdf(G2, g.max_box[None, :])[0] == 7.0


# ---------------------------------------------------------------- gradient / Hessian
def cell_partials(g, x):
    """Independent closed form: slopes of the trilinear cell holding x from its corners."""
    v = g.values.astype(np.float64)
    c = (x - g.min_box) / g.dxyz
    i = np.floor(c).astype(int)
    f = c - i
    c000 = v[i[2], i[1], i[0]]; c100 = v[i[2], i[1], i[0] + 1]
    c010 = v[i[2], i[1] + 1, i[0]]; c110 = v[i[2], i[1] + 1, i[0] + 1]
    c001 = v[i[2] + 1, i[1], i[0]]; c101 = v[i[2] + 1, i[1], i[0] + 1]
    c011 = v[i[2] + 1, i[1] + 1, i[0]]; c111 = v[i[2] + 1, i[1] + 1, i[0] + 1]
    fx, fy, fz = f
    gx = ((1 - fy) * (1 - fz) * (c100 - c000) + fy * (1 - fz) * (c110 - c010) + (1 - fy) * fz * (c101 - c001)
          + fy * fz * (c111 - c011)) / g.dxyz[0]
    gy = ((1 - fx) * (1 - fz) * (c010 - c000) + fx * (1 - fz) * (c110 - c100) + (1 - fx) * fz * (c011 - c001)
          + fx * fz * (c111 - c101)) / g.dxyz[1]
    gz = ((1 - fx) * (1 - fy) * (c001 - c000) + fx * (1 - fy) * (c101 - c100) + (1 - fx) * fy * (c011 - c010)
          + fx * fy * (c111 - c110)) / g.dxyz[2]
    hxy = ((1 - fz) * (c110 - c010 - c100 + c000) + fz * (c111 - c011 - c101 + c001)) / (g.dxyz[0] * g.dxyz[1])
    return np.array([gx, gy, gz]), hxy


def random_grid(seed=1):
    rng = np.random.default_rng(seed)
    vals = rng.normal(size=(10, 11, 12)).astype(np.float32)
    g = W.SDFGrid(vals, np.array([-0.5, -0.4, -0.3]), None, np.array([0.1, 0.09, 0.08]))
    g.max_box = g.min_box + (np.array(g.n) - 1) * g.dxyz
    return g, O.GridRef.from_grid(g)


def test_fd_gradient_equals_cell_partials_away_from_faces():
    g, G = random_grid()
    rng = np.random.default_rng(2)
    x = rand_inside(g, rng, 400, margin=0.01)
    c = (x - g.min_box) / g.dxyz
    fr = c - np.floor(c)
    x = x[np.all((fr > 1e-3) & (fr < 1 - 1e-3), axis=1)]
    gfd, hfd = O.query_dsdf(G, x, O.MODE_REFERENCE)
    gan, han = O.query_dsdf(G, x, O.MODE_ANALYTIC)
    for k in range(len(x)):
        ref_g, ref_hxy = cell_partials(g, x[k])
        assert np.allclose(gfd[k], ref_g, rtol=1e-8, atol=1e-8)
        assert np.allclose(gan[k], ref_g, rtol=1e-12, atol=1e-12)
        assert abs(hfd[k][0, 1] - ref_hxy) < 1e-3 * max(1.0, abs(ref_hxy))
        assert abs(han[k][0, 1] - ref_hxy) < 1e-9 * max(1.0, abs(ref_hxy))
        assert np.all(np.abs(np.diag(hfd[k])) < 1e-3)          # rounding noise of the FD diagonal
        assert np.all(np.diag(han[k]) == 0.0)


def test_fd_blend_and_hessian_spike_at_a_face():
    """x = x_face - delta (delta < eps): FD slope = (sR (h - delta) + sL (h + delta)) / 2h,
    FD curvature = (sR - sL)(h - delta) / h^2."""
    g, G = random_grid(3)
    h = 1e-5
    rng = np.random.default_rng(4)
    for _ in range(20):
        i = rng.integers(2, 9)
        xf = g.min_box[0] + i * g.dxyz[0]
        y, z = rng.uniform(g.min_box[1] + 0.05, g.max_box[1] - 0.05), rng.uniform(g.min_box[2] + 0.05, g.max_box[2] - 0.05)
        delta = rng.uniform(0.1, 0.9) * h
        x = np.array([xf - delta, y, z])
        sL = cell_partials(g, np.array([xf - 0.5 * g.dxyz[0], y, z]))[0][0]
        sR = cell_partials(g, np.array([xf + 0.5 * g.dxyz[0], y, z]))[0][0]
        gfd, hfd = O.query_dsdf(G, x[None, :], O.MODE_REFERENCE)
        expect_g = (sR * (h - delta) + sL * (h + delta)) / (2 * h)
        expect_h = (sR - sL) * (h - delta) / h ** 2
        assert abs(gfd[0][0] - expect_g) < 1e-6 * max(1.0, abs(expect_g))
        assert abs(hfd[0][0, 0] - expect_h) < 1e-4 * max(1.0, abs(expect_h))


# ---------------------------------------------------------------- math::exp
def test_se3_exp_matches_matrix_exponential():
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = rng.normal(size=6) * rng.choice([1e-14, 1e-3, 0.5, 2.0])
        w, v = q[:3], q[3:]
        T = np.zeros((4, 4))
        T[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
        T[:3, 3] = v
        assert np.allclose(O.se3_exp(q), scipy.linalg.expm(T), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- collision derivatives
def test_collision_derivative_blocks_match_central_fd():
    g = W.make_grid(lambda x, y, z: W.sdf_box_cyl_hole(x, y, z, 1.0, 0.52), (64, 64, 64), (-1.575,) * 3, (0.05,) * 3)
    G = O.GridRef.from_grid(g)
    rng = np.random.default_rng(6)
    E2 = W.pose(W.rotation_exp([0.05, -0.02, 0.1]), [0.02, -0.01, 0.03])
    phi2 = rng.normal(0, 0.3, 6)
    xw = rng.uniform(-0.9, 0.9, size=(40, 3))
    xd = rng.normal(size=(40, 3))
    base = O.collision(G, E2, phi2, 0.01, xw, xd, mode=O.MODE_ANALYTIC, derivs=True)
    h = 1e-6

    def vals(E=E2, ph=phi2, x=xw, v=xd):
        c = O.collision(G, E, ph, 0.01, x, v, mode=O.MODE_ANALYTIC)
        return {k: c[k] for k in ("d", "n", "ddot", "tdot", "xi2")}

    # columns in xw, xw_dot, q2 (right-multiplied exp), phi2
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        p, m = vals(x=xw + e), vals(x=xw - e)
        for name, blk in (("d", "dd_dxw"), ("n", "dn_dxw"), ("ddot", "dddot_dxw"), ("tdot", "dtdot_dxw"), ("xi2", "dxi2_dxw")):
            fd = (p[name] - m[name]) / (2 * h)
            an = base[blk][:, k] if base[blk].ndim == 2 else base[blk][:, :, k]
            assert np.allclose(an, fd, rtol=1e-4, atol=1e-4), blk
        p, m = vals(v=xd + e), vals(v=xd - e)
        assert np.allclose(base["dddot_dxwdot"][:, k], (p["ddot"] - m["ddot"]) / (2 * h), atol=1e-6)
        assert np.allclose(base["dtdot_dxwdot"][:, :, k], (p["tdot"] - m["tdot"]) / (2 * h), atol=1e-6)
    for k in range(6):
        dq = np.zeros(6)
        dq[k] = h
        p, m = vals(E=E2 @ O.se3_exp(dq)), vals(E=E2 @ O.se3_exp(-dq))
        for name, blk in (("d", "dd_dq2"), ("n", "dn_dq2"), ("ddot", "dddot_dq2"), ("tdot", "dtdot_dq2"), ("xi2", "dxi2_dq2")):
            fd = (p[name] - m[name]) / (2 * h)
            an = base[blk][:, k] if base[blk].ndim == 2 else base[blk][:, :, k]
            assert np.allclose(an, fd, rtol=1e-4, atol=1e-4), blk
        p, m = vals(ph=phi2 + dq), vals(ph=phi2 - dq)
        assert np.allclose(base["dddot_dphi2"][:, k], (p["ddot"] - m["ddot"]) / (2 * h), atol=1e-6)
        assert np.allclose(base["dtdot_dphi2"][:, :, k], (p["tdot"] - m["tdot"]) / (2 * h), atol=1e-6)


# ---------------------------------------------------------------- force
def plane_grid():
    g = W.make_grid(lambda x, y, z: z + 0 * x, (21, 21, 21), (-1.0, -1.0, -1.0), (0.1, 0.1, 0.1))
    return g, O.GridRef.from_grid(g)


def test_single_contact_on_a_plane_by_hand():
    """phi = z, body 2 at rest at the origin, one contact at xi = (0.1, 0.2, -0.01):
    d = -0.01, n = +z, contact force kn*0.01 along +z on body 1 at xi, the opposite
    force on body 2 at xi2 = (0.1, 0.2, 0); with velocity v1 = (0.5, 0, -1): ddot = -1,
    damping adds c*ddot*d, static friction -kt*(0.5, 0, 0) on body 1."""
    g, G = plane_grid()
    kn, kt, mu, c = 1e4, 100.0, 0.5, 10.0
    xi = np.array([[0.1, 0.2, -0.01]])
    d = -0.01
    # frictionless, at rest
    r = O.force_fixed_contacts(G, xi, np.eye(4), np.zeros(6), np.eye(4), np.zeros(6), 0.0, kn, 0.0, 0.0, 0.0)
    f1 = np.array([0, 0, -kn * d])
    assert np.allclose(r["fm"][3:6], f1)
    assert np.allclose(r["fm"][0:3], np.cross(xi[0], f1))
    xi2 = np.array([0.1, 0.2, 0.0])
    assert np.allclose(r["fm"][9:12], -f1)
    assert np.allclose(r["fm"][6:9], np.cross(xi2, -f1))
    # moving: v1 = (0.5, 0, -1) in body 1 frame (= world), damping and static friction
    phi1 = np.array([0, 0, 0, 0.5, 0.0, -1.0])
    r = O.force_fixed_contacts(G, xi, np.eye(4), phi1, np.eye(4), np.zeros(6), 0.0, kn, kt, mu, c)
    lam = -kn * d + c * (-1.0) * d
    fn = np.array([0, 0, lam])
    tdot = np.array([0.5, 0, 0])
    assert mu * np.linalg.norm(np.concatenate([np.cross(xi[0], fn), fn])) >= kt * 0.5   # static branch
    f1 = fn - kt * tdot
    assert np.allclose(r["fm"][3:6], f1)
    assert np.allclose(r["fm"][0:3], np.cross(xi[0], fn) - kt * np.cross(xi[0], tdot))
    assert np.allclose(r["fm"][9:12], -fn + kt * tdot)
    # dynamic branch: large kt
    r = O.force_fixed_contacts(G, xi, np.eye(4), phi1, np.eye(4), np.zeros(6), 0.0, kn, 1e6, mu, c)
    fcn = np.linalg.norm(np.concatenate([np.cross(xi[0], fn), fn]))
    assert np.allclose(r["fm"][3:6], fn - mu * fcn * np.array([1.0, 0, 0]))


@pytest.fixture(scope="module")
def peg():
    w = W.peg_in_hole(n_poses=12)
    return w, O.GridRef.from_grid(w.grids[0]), W.round_points_f32(w.point_sets[0])


def _eval(G, pts, ins, **kw):
    p = dict(kn=ins.kn, kt=ins.kt, mu=ins.mu, damping=ins.damping)
    p.update(kw)
    oi = O.make_instance(0, pts, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, p["kn"], p["kt"], p["mu"],
                         p["damping"], kw.get("scale", 1.0), kw.get("contact_scale", 1.0))
    return O.eval_instance([G], oi, O.FLAG_DERIV | O.FLAG_P6)


def test_world_force_and_torque_balance(peg):
    w, G, pts = peg
    for ins in w.instances:
        R1, p1 = ins.E1[:3, :3], ins.E1[:3, 3]
        R2, p2 = ins.E2[:3, :3], ins.E2[:3, 3]
        r = _eval(G, pts, ins)
        if not r["n_active"]:
            continue
        f1w, f2w = R1 @ r["fm"][3:6], R2 @ r["fm"][9:12]
        assert np.linalg.norm(f1w + f2w) <= 1e-12 * r["scale_fm"]
        rn = _eval(G, pts, ins, mu=0.0)         # the normal pair only: torques balance too
        t1w = R1 @ rn["fm"][0:3] + np.cross(p1, R1 @ rn["fm"][3:6])
        t2w = R2 @ rn["fm"][6:9] + np.cross(p2, R2 @ rn["fm"][9:12])
        assert np.linalg.norm(t1w + t2w) <= 1e-11 * rn["scale_fm"]


def test_kn_linearity_and_zero_kt(peg):
    w, G, pts = peg
    for ins in w.instances[:6]:
        a = _eval(G, pts, ins, mu=0.0, damping=0.0)
        b = _eval(G, pts, ins, mu=0.0, damping=0.0, kn=2 * ins.kn)
        assert np.allclose(b["fm"], 2 * a["fm"], rtol=1e-12, atol=1e-9)
        c = _eval(G, pts, ins, kt=0.0)
        d = _eval(G, pts, ins, mu=0.0)
        assert np.allclose(c["fm"], d["fm"], rtol=1e-12, atol=1e-9)


def test_km_dm_match_central_fd(peg):
    w, G, pts = peg
    h = 1e-7
    for ins in w.instances[:4]:
        r = _eval(G, pts, ins)
        if not r["n_active"]:
            continue
        xi = pts[r["active"]]

        def fm(E1=ins.E1, E2=ins.E2, phi1=ins.phi1, phi2=ins.phi2):
            return O.force_fixed_contacts(G, xi, E1, phi1, E2, phi2, ins.col_th, ins.kn, ins.kt, ins.mu,
                                          ins.damping)["fm"]
        K = np.zeros((12, 12))
        D = np.zeros((12, 12))
        for k in range(6):
            dq = np.zeros(6)
            dq[k] = h
            K[:, k] = (fm(E1=ins.E1 @ O.se3_exp(dq)) - fm(E1=ins.E1 @ O.se3_exp(-dq))) / (2 * h)
            K[:, 6 + k] = (fm(E2=ins.E2 @ O.se3_exp(dq)) - fm(E2=ins.E2 @ O.se3_exp(-dq))) / (2 * h)
            D[:, k] = (fm(phi1=ins.phi1 + dq) - fm(phi1=ins.phi1 - dq)) / (2 * h)
            D[:, 6 + k] = (fm(phi2=ins.phi2 + dq) - fm(phi2=ins.phi2 - dq)) / (2 * h)
        assert np.linalg.norm(r["Km"] - K) <= 1e-5 * r["scale_Km"]
        assert np.linalg.norm(r["Dm"] - D) <= 1e-5 * r["scale_Dm"]
        err, rel = O.print_error_metric(r["Km"], K)
        assert not (rel > 1e-5 and err > 1e-5 * r["scale_Km"])


def test_p6_column(peg):
    """d fm / d contact_scale: equals FD when damping = 0; the reference's dynamic-friction
    term (fd / contact_scale) is exact only then, and that formula is what is kept."""
    w, G, pts = peg
    for damping, exact in ((0.0, True), (10.0, False)):
        for ins in w.instances[:4]:
            r = _eval(G, pts, ins, damping=damping)
            if not r["n_active"]:
                continue
            xi = pts[r["active"]]

            def fm(cs):
                return O.force_fixed_contacts(G, xi, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt,
                                              ins.mu, damping, 1.0, cs)["fm"]
            fd = (fm(1 + 1e-6) - fm(1 - 1e-6)) / 2e-6
            if exact:
                assert np.linalg.norm(r["dfm_dp6"] - fd) <= 1e-6 * r["scale_fm"]
            else:
                # same normal and static-friction terms; only the dynamic-friction term differs
                rn = _eval(G, pts, ins, damping=damping, mu=0.0)
                fdn = (O.force_fixed_contacts(G, xi, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt,
                                              0.0, damping, 1.0, 1 + 1e-6)["fm"]
                       - O.force_fixed_contacts(G, xi, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt,
                                                0.0, damping, 1.0, 1 - 1e-6)["fm"]) / 2e-6
                assert np.linalg.norm(rn["dfm_dp6"] - fdn) <= 1e-6 * r["scale_fm"]

"""The C++ drop-in classes (atac._redmax: Body, BodySDFObj, ForceGeneralSDFContact)
on the GPU against the FP64 oracle: the public wrappers' accumulation into global
fm / Km / Dm at each body's _index (ForceGeneralSDFContact.cpp:46-144), the p6
design column, BodySDFObj::distance / collision, the contact history and the
finite-difference self-checks."""
import numpy as np
import pytest

from atac import workloads as W
from helpers import rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def scene(gpu):
    from atac import _redmax as R
    w = W.peg_in_hole(n_poses=12)
    g = w.grids[0]
    pts = W.round_points_f32(w.point_sets[0])
    peg = R.Body("peg", pts, 0)
    hole = R.BodySDFObj("hole", g.values.astype(np.float64), g.min_box, g.max_box, g.dxyz, 0.01, np.zeros((1, 3)), 6)
    f = R.ForceGeneralSDFContact(peg, hole, W.KN, W.KT, W.MU, W.DAMPING)
    return R, w, pts, peg, hole, f


def _oracle(w, pts, ins, flags=O.FLAG_DERIV | O.FLAG_P6):
    G = O.GridRef.from_grid(w.grids[0])
    return O.eval_instance([G], O.make_instance(0, pts, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn,
                                                ins.kt, ins.mu, ins.damping), flags)


def test_force_with_derivative_matches_oracle(scene):
    R, w, pts, peg, hole, f = scene
    print("[case] redmax ForceGeneralSDFContact peg-in-hole 12 poses", flush=True)
    for ins in w.instances:
        peg.E_0i, peg.phi = ins.E1, ins.phi1
        hole.E_0i, hole.phi = ins.E2, ins.phi2
        fm, Km, Dm = f.computeForceWithDerivative(12)
        r = _oracle(w, pts, ins)
        assert f.last_active_count() == r["n_active"]
        if r["n_active"]:
            assert rel_err(fm, r["fm"], r["scale_fm"]) <= TOL
            assert rel_err(Km, r["Km"], r["scale_Km"]) <= TOL
            assert rel_err(Dm, r["Dm"], r["scale_Dm"]) <= TOL
        # the wrench alone: the same contacts and summation order; the two evaluations are
        # separate kernels whose FP32 contact arithmetic may round differently (the compiler
        # fuses multiply-adds per kernel), so they agree to FP32 rounding of the per-contact scale
        fm2 = f.computeForce(12)
        assert rel_err(fm2, fm, max(1.0, r["scale_fm"])) <= 1e-7


def test_scatter_into_global_arrays(scene):
    """Bodies at other indices of a larger system: blocks land at _index and accumulate."""
    R, w, pts, peg, hole, f = scene
    ins = w.instances[1]
    peg2 = R.Body("peg2", pts, 18)
    hole2 = R.BodySDFObj("hole2", w.grids[0].values.astype(np.float64), w.grids[0].min_box, w.grids[0].max_box,
                         w.grids[0].dxyz, 0.01, np.zeros((1, 3)), 6)
    f2 = R.ForceGeneralSDFContact(peg2, hole2, W.KN, W.KT, W.MU, W.DAMPING)
    peg2.E_0i, peg2.phi, hole2.E_0i, hole2.phi = ins.E1, ins.phi1, ins.E2, ins.phi2
    fm, Km, Dm = f2.computeForceWithDerivative(30)
    r = _oracle(w, pts, ins)
    idx = np.r_[18:24, 6:12]
    assert rel_err(fm[idx], r["fm"], r["scale_fm"]) <= TOL
    assert rel_err(Km[np.ix_(idx, idx)], r["Km"], r["scale_Km"]) <= TOL
    rest = np.ones(30, dtype=bool)
    rest[idx] = False
    assert np.all(fm[rest] == 0.0) and np.all(Km[rest] == 0.0)
    assert "hole2" in peg2.get_contact_bodies() and "peg2" in hole2.get_contact_bodies()


def test_design_column(scene):
    R, w, pts, peg, hole, f = scene
    ins = w.instances[2]
    peg.E_0i, peg.phi, hole.E_0i, hole.phi = ins.E1, ins.phi1, ins.E2, ins.phi2
    peg.p6_column = 3
    try:
        fm, Km, Dm, dp = f.computeForceWithDesignDerivative(12, 5)
    finally:
        peg.p6_column = -1
    r = _oracle(w, pts, ins)
    assert rel_err(dp[:, 3], r["dfm_dp6"], r["scale_fm"]) <= TOL
    assert np.all(dp[:, [0, 1, 2, 4]] == 0.0)


def test_body_sdf_queries(scene):
    R, w, pts, peg, hole, f = scene
    E2 = W.pose(W.rotation_exp([0.02, -0.03, 0.05]), [0.01, 0.0, -0.02])
    phi2 = np.array([0.1, -0.2, 0.3, 0.05, 0.0, -0.1])
    hole.E_0i, hole.phi = E2, phi2
    G = O.GridRef.from_grid(w.grids[0])
    rng = np.random.default_rng(1)
    for _ in range(10):
        xw = rng.uniform(-0.9, 0.9, 3)
        xd = rng.normal(size=3)
        d = hole.distance(xw)
        assert abs(d - O.distance(G, E2, 0.01, xw[None])[0]) < 1e-6
        c = hole.collision(xw, xd, True)
        ref = O.collision(G, E2, phi2, 0.01, xw[None], xd[None], derivs=True)
        assert abs(c["d"] - ref["d"][0]) < 1e-9
        assert np.allclose(c["n"].ravel(), ref["n"][0], atol=1e-7)
        assert np.allclose(c["dn_dq2"], ref["dn_dq2"][0], atol=1e-5 * max(1.0, np.abs(ref["dn_dq2"]).max()))
        assert np.allclose(c["dtdot_dxw"], ref["dtdot_dxw"][0], atol=1e-5 * max(1.0, np.abs(ref["dtdot_dxw"]).max()))


def test_fd_self_checks(scene):
    R, w, pts, peg, hole, f = scene
    ins = w.instances[1]
    peg.E_0i, peg.phi, hole.E_0i, hole.phi = ins.E1, ins.phi1, ins.E2, ins.phi2
    assert f.test_derivatives() < 1e-4
    assert f.test_derivatives_runtime() < 1e-4

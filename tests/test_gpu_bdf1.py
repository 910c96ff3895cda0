"""Batched BDF1 stepping of parts on translational joints (atac._redmax.BatchedBDF1,
Simulation::integration_BDF1 / newton with the SDF contact forces evaluated on the
GPU) against the FP64 CPU reference of the same integrator driven by the oracle's
forces; free flight against the BDF1 recursion in closed form."""
import numpy as np
import pytest

from atac import workloads as W
from bdf1_reference import ReferenceBDF1

pytestmark = pytest.mark.gpu

KN, KT, MU, DAMP, COL = 1e4, 1e3, 0.3, 10.0, 0.01


@pytest.fixture(scope="module")
def scene(gpu):
    from atac import _redmax as R
    w = W.peg_in_hole(n_points=2000, n_poses=8)
    g = w.grids[0]
    hole = R.BodySDFObj("hole", g.values.astype(np.float64), g.min_box, g.max_box, g.dxyz, COL, np.zeros((1, 3)), 0)
    ins = w.instances[3]                     # a pose in contact
    E_peg = ins.E1.copy()
    parts_gpu = [(w.point_sets[0], None, E_peg), (np.zeros((0, 3)), hole, np.eye(4))]
    parts_ref = [(w.point_sets[0], None, E_peg), (np.zeros((0, 3)), g, np.eye(4))]
    return R, w, hole, parts_gpu, parts_ref


def test_free_flight_closed_form(scene):
    R, w, hole, parts_gpu, _ = scene
    h, m = 1e-3, 2.0
    grav = np.array([0.0, 0.0, -9.8])
    sim = R.BatchedBDF1(parts_gpu, [], h=h, gravity=grav)
    u = np.array([3.0, -1.0, 0.5])
    c = sim.add_candidate(0, m, np.zeros(3), np.array([0.1, 0.0, 0.2]), u)
    n = 25
    sim.step(n)
    a = u / m + grav
    # BDF1: qd_{k+1} = qd_k + h a, q_{k+1} = q_k + h qd_{k+1}
    qd = np.array([0.1, 0.0, 0.2]) + n * h * a
    q = n * h * np.array([0.1, 0.0, 0.2]) + h * h * a * n * (n + 1) / 2
    assert np.allclose(np.ravel(sim.q(c)), q, rtol=1e-12, atol=1e-14)
    assert np.allclose(np.ravel(sim.qdot(c)), qd, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("precise", [False, True])
def test_contact_steps_match_fp64_reference(scene, precise):
    R, w, hole, parts_gpu, parts_ref = scene
    h, m, n = 1e-3, 1.5, 12
    pushes = [np.array([400.0, 0, 0]), np.array([0, 0, -300.0]), np.array([-200.0, 250.0, 0]), np.zeros(3)]
    sim = R.BatchedBDF1(parts_gpu, [(0, 1)], kn=KN, kt=KT, mu=MU, damping=DAMP, col_th=COL, h=h, precise=precise)
    ids = [sim.add_candidate(0, m, np.zeros(3), np.zeros(3), u) for u in pushes]
    ref = ReferenceBDF1(parts_ref, [(0, 1)], KN, KT, MU, DAMP, COL, h, np.zeros(3))
    for s in range(n):
        sim.step(1)
    dev = 0.0
    for c, u in zip(ids, pushes):
        q_ref, qd_ref, info = ref.step(0, m, np.zeros(3), np.zeros(3), u, n)
        q = np.ravel(sim.q(c))
        scale = max(np.linalg.norm(q_ref), 1e-4)
        assert np.linalg.norm(q - q_ref) <= 1e-4 * scale, (c, q, q_ref)
        assert sum(ok for ok, _ in info) >= n - 1
        q_free = h * h * (u / m) * n * (n + 1) / 2      # the same push without contact
        dev = max(dev, np.linalg.norm(q_ref - q_free))
    assert dev > 1e-5, "the contact force should change the motion"
    assert sim.n_evaluations > 0


def test_contact_changes_the_motion(scene):
    """Pushing the peg sideways into the hole wall is resisted: it moves less than in
    free flight."""
    R, w, hole, parts_gpu, _ = scene
    h, m, n = 1e-3, 1.5, 20
    u = np.array([2000.0, 0, 0])
    free = R.BatchedBDF1(parts_gpu, [], h=h)
    cont = R.BatchedBDF1(parts_gpu, [(0, 1)], kn=1e6, kt=KT, mu=0.0, damping=0.0, col_th=COL, h=h)
    a = free.add_candidate(0, m, np.zeros(3), np.zeros(3), u)
    b = cont.add_candidate(0, m, np.zeros(3), np.zeros(3), u)
    free.step(n)
    cont.step(n)
    assert np.ravel(cont.q(b))[0] < np.ravel(free.q(a))[0]

"""Workload 5, the multi-part step loop: a fixed base plate and free parts (superquadrics
and square-hole blocks on free3d-exp joints) resting on it under gravity, every ordered
pair a general_SDF_contact, the AABB broad phase choosing the pairs each evaluation
runs, BDF1 with the FP64 Newton solve on the host (Simulation.cpp:999-1098, 1248-1270).

A reduced scene (4 parts, about 1,500 contact points each) is stepped against the FP64
reference integrator over all 12 pairs; the full-size scene (20 parts, about 100k
contact points and 253^3 grid nodes each) is checked by its Newton residuals and by
the broad phase giving exactly the forces of evaluating every pair."""
import numpy as np
import pytest

from atac import _redmax as R
from atac import workloads as W
from sim_reference import RefScene

pytestmark = pytest.mark.gpu

KN, KT, MU, DAMP = 1e4, 1e3, 0.3, 10.0


def test_reduced_scene_matches_reference_integrator(gpu, tmp_path):
    xml, names = W.multi_part_scene(str(tmp_path), n_parts=4, points_per_part=1500, grid_res=40, kn=KN, kt=KT, mu=MU,
                                    damping=DAMP)
    sim = R.Simulation(xml, str(tmp_path), precise=True)
    assert sim.ndof_r == 18
    n = len(names)
    ref = RefScene(sim, [(names[0], "fixed", 0.0)] + [(m, "free3d-exp", 0.0) for m in names[1:]],
                   [(a, b, KN, KT, MU, DAMP) for a in range(n) for b in range(n) if a != b], [0, 0, -9.8],
                   sim.timestep)
    q0 = np.zeros(sim.ndof_r)
    ref.set_E_pj0(sim, q0)
    sim.reset()
    steps = 6
    sim.forward(steps)
    assert sim.is_converged()
    st = sim.stats()
    qh, qdh = sim.get_q_his(), sim.get_qdot_his()
    q, qd = q0.copy(), q0.copy()
    errs, res = [], []
    for k in range(steps):
        q, qd, r = ref.step(q, qd)
        assert r < 1e-9
        errs.append(np.abs(qh[k + 1] - q).max())
        res.append(np.linalg.norm(ref.residual(qh[k + 1], qh[k], qdh[k])))
    disp = np.abs(qh[-1]).max()
    print(f"  reduced scene: max |q - q_ref| {max(errs):.3g} (displacement {disp:.3g}), residual {max(res):.3g}; "
          f"{st}", flush=True)
    assert st["pairs_culled"] > 0 and st["pairs_evaluated"] > 0
    assert all(sorted(sim.get_contact_bodies(m)) == [names[0]] for m in names[1:])
    assert max(res) <= 2e-8 and max(errs) <= 2e-6


def test_full_size_scene(gpu, tmp_path):
    xml, names = W.multi_part_scene(str(tmp_path), kn=KN, kt=KT, mu=MU, damping=DAMP)
    sim = R.Simulation(xml, str(tmp_path))
    assert len(names) == 20 and sim.ndof_r == 19 * 6
    n_points = sum(len(sim.get_body_contact_points(m)) for m in names)
    shapes = [sim.get_body_sdf(m)["values"].shape for m in names[:3]]
    print(f"  full size: {n_points} contact points, grids {shapes} ...", flush=True)
    assert 1.8e6 <= n_points <= 2.2e6
    assert all(min(s) >= 250 for s in shapes)
    sim.reset()
    steps = 3
    sim.forward(steps)
    assert sim.is_converged()
    st = sim.stats()
    print(f"  {st}", flush=True)
    assert st["pairs_culled"] > st["pairs_evaluated"] > 0
    # each step solves its BDF1 equation to the Newton tolerance
    qh, qdh = sim.get_q_his(), sim.get_qdot_his()
    h = sim.timestep
    for k in range(steps):
        sim.set_state(qh[k + 1], (qh[k + 1] - qh[k]) / h)
        o = sim.computeMatrices(False)
        g = o["M"] @ (qh[k + 1] - qh[k] - h * qdh[k]) - h * h * o["f"]
        assert np.linalg.norm(g) < 1e-8
    # the broad phase only leaves out pairs that contribute nothing: the same f, K, D
    sim.set_state(qh[-1], qdh[-1])
    on = sim.computeMatrices(True)
    sim.set_broad_phase(False)
    off = sim.computeMatrices(True)
    sim.set_broad_phase(True)
    for key in ("M", "f", "K", "D"):
        np.testing.assert_array_equal(on[key], off[key], err_msg=key)

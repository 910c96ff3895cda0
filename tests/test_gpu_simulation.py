"""The simulator subset with SDF contact on the GPU (redmax_py.Simulation's interface,
python_interface.cpp:84-276): scenes built from the reference's XML format, stepped by
BDF1 with the contact forces of every Newton iteration evaluated in one batch, against
the FP64 reference integrator (tests/sim_reference.py) over the same contacts; the
distance, contact-history, contact-scale and .sdf-cache methods against the oracle."""
import os

import numpy as np
import pytest

import scene_meshes as SM
from atac import _redmax as R
from atac import workloads as W
from oracle import oracle as O
from sim_reference import RefScene

pytestmark = pytest.mark.gpu

KN, KT, MU, DAMP = 1e3, 1e2, 0.3, 1.0


def _link(name, joint, filename, jpos="0 0 0", jquat="1 0 0 0", extra=""):
    return f'''<robot><link name="{name}">
      <joint name="{name}" type="{joint}" axis="0. 0. 0." pos="{jpos}" quat="{jquat}" frame="WORLD" damping="0"/>
      <body name="{name}" type="SDF" filename="{filename}" pos="0 0 0" quat="1 0 0 0" scale="1 1 1"
            transform_type="OBJ_TO_JOINT" density="1" dx="0.05" col_th="0.01" mu="0" {extra}/></link></robot>'''


def _scene(moving_joint, gravity="0. 0. -9.8", moving_extra="", jpos="0 0 0.287", jquat="0.999 0.03 -0.02 0.0"):
    return f'''<redmax model="box_on_ground">
<option integrator="BDF1" timestep="1e-3" gravity="{gravity}"/>
<default><general_SDF_contact kn="{KN}" kt="{KT}" mu="{MU}" damping="{DAMP}"/></default>
{_link("box", moving_joint, "box.obj", jpos=jpos, jquat=jquat, extra=moving_extra)}
{_link("ground", "fixed", "ground.obj", jpos="0 0 -0.5")}
<contact>
  <general_SDF_contact general_body="box" SDF_body="ground"/>
  <general_SDF_contact general_body="ground" SDF_body="box"/>
</contact>
</redmax>'''


@pytest.fixture(scope="module")
def assets(gpu, tmp_path_factory):
    d = tmp_path_factory.mktemp("scene")
    SM.write_obj(os.path.join(d, "box.obj"), *SM.box_mesh([1.0, 0.8, 0.6], 6))
    SM.write_obj(os.path.join(d, "ground.obj"), *SM.box_mesh([4.0, 4.0, 1.0], 8))
    return str(d)


def _reference(sim, joint_type, q0, gravity=(0, 0, -9.8)):
    ref = RefScene(sim, [("box", joint_type, 0.0), ("ground", "fixed", 0.0)],
                   [(0, 1, KN, KT, MU, DAMP), (1, 0, KN, KT, MU, DAMP)], gravity, sim.timestep)
    ref.set_E_pj0(sim, np.concatenate([q0, []]))
    return ref


def _run(sim, ref, q0, qd0, n, ext=None):
    """Steps both; returns the trajectory differences, the largest residual of the
    product's steps in the reference's BDF1 equation, and the reference's last state."""
    sim.set_state_init(q0, qd0)
    sim.reset()
    if ext is not None:
        sim.set_body_external_force("box", ext)
        ref.bodies[0].f_ext = np.asarray(ext, dtype=np.float64)
    sim.forward(n)
    assert sim.is_converged()
    q, qd = q0.copy(), qd0.copy()
    errs, res_prod = [], []
    qh, qdh = sim.get_q_his(), sim.get_qdot_his()
    for k in range(n):
        q, qd, res = ref.step(q, qd)
        assert res < 1e-9
        errs.append(np.abs(qh[k + 1] - q).max())
        res_prod.append(np.linalg.norm(ref.residual(qh[k + 1], qh[k], qdh[k])))
    return np.array(errs), max(res_prod), q


# The product's Newton stops at |g| < 1e-8 (Simulation.cpp:1007), so its steps satisfy
# the reference's equation to that tolerance (the FP32 contact forces add h^2 |df|, far
# below it), and its states differ from the reference's by up to |H^-1| 1e-8: with
# rotational inertias near 0.02 a few 1e-7.
RES_TOL = 2e-8


@pytest.mark.parametrize("precise", [True, False])
def test_free_box_on_ground_matches_reference_integrator(assets, precise):
    sim = R.Simulation(_scene("free3d-exp"), assets, precise=precise)
    assert sim.ndof_r == 6
    q0 = np.zeros(6)
    qd0 = np.array([0.02, -0.01, -0.05, 0.1, -0.2, 0.05])
    ref = _reference(sim, "free3d-exp", q0)
    errs, res, q = _run(sim, ref, q0, qd0, 12)
    st = sim.stats()
    disp = np.abs(q - q0).max()
    print(f"  precise={precise}: max |q - q_ref| {errs.max():.3g} (displacement {disp:.3g}), product steps' "
          f"residual in the reference equation {res:.3g}; {st}", flush=True)
    assert st["pairs_evaluated"] > 0 and st["evaluations"] > 0
    assert sorted(sim.get_contact_bodies("box")) == ["ground"]
    assert res <= RES_TOL
    assert errs.max() <= 2e-6


def test_translational_push_matches_reference_integrator(assets):
    """The disassembly planner's setting: the moving part on a translational joint, pushed by
    an external force (run_joint_plan.py:161, 263-272), against the fixed part."""
    sim = R.Simulation(_scene("translational", gravity="0. 0. 1e-12", jquat="1 0 0 0"), assets, precise=True)
    q0, qd0 = np.zeros(3), np.zeros(3)
    ref = _reference(sim, "translational", q0, gravity=(0, 0, 1e-12))
    ext = np.array([0, 0, 0, 30.0, -10.0, -5.0])
    errs, res, q = _run(sim, ref, q0, qd0, 15, ext=ext)
    print(f"  translational: max |q - q_ref| {errs.max():.3g}, residual {res:.3g}, q {q}", flush=True)
    assert res <= RES_TOL and errs.max() <= 2e-6


def test_distance_contact_scale_and_history(assets):
    sim = R.Simulation(_scene("free3d-exp"), assets, precise=True)
    # get_body_distance: min over the world vertices of the other body's SDF distance
    for a, b in (("box", "ground"), ("ground", "box")):
        g = sim.get_body_sdf(b)
        G = O.GridRef(g["values"], g["min_box"].ravel(), g["max_box"].ravel(), g["dxyz"])
        xw = sim.get_body_vertices(a, world_frame=True).T
        d = O.distance(G, sim.get_body_E0i(b), g["col_th"], xw)
        assert abs(sim.get_body_distance(a, b) - d.min()) <= 1e-6 * max(1.0, abs(d.min()))
    assert sim.body_in_contact("box", "ground")
    # the contact scale multiplies kn and kt (ForceGeneralSDFContact.cpp:155-156)
    sim.set_contact_scale(0.5)
    q0, qd0 = np.zeros(6), np.array([0.0, 0.0, -0.05, 0.0, 0.0, 0.0])
    ref = _reference(sim, "free3d-exp", q0)
    ref.scale = 0.5
    errs, res, _ = _run(sim, ref, q0, qd0, 5)
    assert res <= RES_TOL and errs.max() <= 2e-6
    sim.clear_contact_bodies("box")
    assert sim.get_contact_bodies("box") == []


def test_broad_phase_skips_pairs_that_cannot_touch(assets):
    """A box 2 m above the ground: every pair is culled, no contact evaluation runs, and
    the motion is free fall."""
    sim = R.Simulation(_scene("free3d-exp", jpos="0 0 2.5", jquat="1 0 0 0"), assets)
    sim.reset()
    sim.forward(5)
    st = sim.stats()
    assert st["evaluations"] == 0 and st["pairs_culled"] > 0 and st["pairs_evaluated"] == 0
    assert not sim.body_in_contact("box", "ground")
    h = sim.timestep
    z = 0.0
    vz = 0.0
    for _ in range(5):     # BDF1 free fall: v1 = v0 + h g, z1 = z0 + h v1
        vz += -9.8 * h
        z += h * vz
    np.testing.assert_allclose(sim.get_q()[:3], [0, 0, z], atol=1e-12)


def test_sdf_cache_round_trip(assets):
    xml = _scene("free3d-exp", moving_extra='save_sdf="true"')
    sim = R.Simulation(xml, assets)
    path = os.path.join(assets, "box.sdf")
    assert os.path.exists(path)
    g1 = sim.get_body_sdf("box")
    sim2 = R.Simulation(_scene("free3d-exp", moving_extra='load_sdf="true"'), assets)
    g2 = sim2.get_body_sdf("box")
    # the .sdf text keeps 6 significant digits (the stream default, as BodySDFObj::save_SDF
    # writes it, BodySDFObj.cpp:105-116)
    six = np.vectorize(lambda v: float(f"{v:g}"))
    np.testing.assert_array_equal(g2["values"], six(g1["values"].astype(np.float64)).astype(np.float32))
    np.testing.assert_array_equal(g2["min_box"], six(g1["min_box"]))
    sim2.clear_all_saved_sdfs()
    assert not os.path.exists(path)


def test_body_sdf_obj_reference_constructor_and_methods(assets):
    """BodySDFObj(filename, R, p, dx, res, col_th, transform_type, density, ...)
    (BodySDFObj.h:18-27): the mesh body in its principal frame, its grid built from the
    body-frame mesh; surface_contact_pt (BodySDFObj.cpp:215-237) against the oracle's
    collision (xi2 = x - d e) and on a flat field; the collision-derivative self-test;
    clear_saved_SDF."""
    path = os.path.join(assets, "box.obj")
    p = np.array([0.1, 0.2, 0.3])
    b = R.BodySDFObj(path, np.eye(3), p, 0.05, 20, 0.01, transform_type="OBJ_TO_JOINT", density=2.0, save_sdf=True)
    m = 1.0 * 0.8 * 0.6 * 2.0
    assert abs(b.mass - m) < 1e-12
    np.testing.assert_allclose(b.inertia.ravel()[:3], m / 12 * np.array([0.64 + 0.36, 1.0 + 0.36, 1.0 + 0.64]),
                               rtol=1e-12)
    # OBJ_TO_JOINT with R = I: E_ji = [R_oi, p_oi + p]
    np.testing.assert_allclose(b.E_ji[:3, :3], b.E_oi[:3, :3], atol=1e-15)
    np.testing.assert_allclose(b.E_ji[:3, 3], b.E_oi[:3, 3] + p, atol=1e-15)
    g = R.BodySDFObj.from_mesh("same", b.get_vertices(), b.get_faces().T.copy(), 0.05, 20, 0.01, np.zeros((1, 3)))
    np.testing.assert_array_equal(b.values(), g.values())
    np.testing.assert_array_equal(b.min_box, g.min_box)
    assert os.path.exists(os.path.join(assets, "box.sdf"))
    b.clear_saved_SDF()
    assert not os.path.exists(os.path.join(assets, "box.sdf"))
    # surface_contact_pt at points near the surface, in a posed body
    rng = np.random.default_rng(3)
    E = W.pose(W.random_rotation(rng), rng.normal(size=3))
    b.E_0i = E
    G = O.GridRef(b.values().astype(np.float32), b.min_box.ravel(), b.max_box.ravel(), b.dxyz)
    xb = rng.uniform(-0.55, 0.55, size=(20, 3)) * np.array([1.0, 0.8, 0.6])
    xw = xb @ E[:3, :3].T + E[:3, 3]
    ref = O.collision(G, E, np.zeros(6), 0.01, xw, np.zeros_like(xw))
    for k in range(len(xw)):
        np.testing.assert_allclose(b.surface_contact_pt(xw[k]).ravel(), ref["xi2"][k], atol=1e-12)
    # a flat field: no gradient, so the contact point is the query point itself
    flat = R.BodySDFObj("flat", -np.ones((4, 4, 4)), [0, 0, 0], [3, 3, 3], [1, 1, 1], 0.0, np.zeros((1, 3)))
    np.testing.assert_allclose(flat.surface_contact_pt([1.2, 1.7, 0.4]).ravel(), [1.2, 1.7, 0.4], atol=1e-15)
    # the collision-derivative self-test (random pose, twist and interior point)
    err = b.test_collision_derivatives()
    print(f"  test_collision_derivatives: worst relative error {err:.3g}", flush=True)
    assert err < 1e-5


def test_candidates_in_lockstep_match_single_runs(assets):
    """The planner's candidate actions (run_joint_plan.py:593-606: a force or a torque
    along each axis) stepped together: every candidate's trajectory equals, bit for bit,
    the one reset + forward gives it alone; the rounds evaluate all candidates' contact
    pairs in one batch each."""
    sim = R.Simulation(_scene("free3d-exp"), assets)
    names = sim.body_names()
    K = 6
    q0 = np.zeros((K, 6))
    qd0 = np.zeros((K, 6))
    forces = np.zeros((K, len(names), 6))
    for c in range(K):
        forces[c, names.index("box"), c] = 3.0 if c < 3 else 0.0        # torques
        forces[c, names.index("box"), c] += 20.0 if c >= 3 else 0.0     # forces
    sim.reset_stats()
    out = sim.forward_candidates(q0, qd0, forces, 8)
    batched = sim.stats()
    single_evals = 0
    for c in range(K):
        sim.set_state_init(q0[c], qd0[c])
        sim.reset()
        sim.set_body_external_force("box", forces[c, names.index("box")])
        sim.reset_stats()
        sim.forward(8)
        single_evals += sim.stats()["evaluations"]
        np.testing.assert_array_equal(out["q"][c], sim.get_q())
        np.testing.assert_array_equal(out["qdot"][c], sim.get_qdot())
        assert out["converged"][c] == int(sim.is_converged())
    print(f"  {K} candidates x 8 steps: {batched['evaluations']} batched evaluations "
          f"({single_evals} one at a time)", flush=True)
    assert batched["evaluations"] < single_evals


def test_contact_vjp_of_the_last_evaluation(assets):
    """Simulation.contact_vjp, the adjoint of the step's last contact evaluation, against
    that evaluation's own Jacobians (g K for the poses, g D for the velocities) and, for
    kn and kt, against a central difference in the contact scale (which multiplies both,
    ForceGeneralSDFContact.cpp:155-156, so d fm/d scale = kn d fm/d kn + kt d fm/d kt)."""
    sim = R.Simulation(_scene("free3d-exp"), assets, precise=True)
    with pytest.raises(Exception):
        sim.contact_vjp(np.zeros((2, 12)))
    sim.set_state_init(np.zeros(6), np.array([0.02, -0.01, -0.05, 0.1, -0.2, 0.05]))
    sim.reset()
    sim.forward(4)
    sim.computeMatrices(True)
    e = sim.last_contact_eval()
    P = len(e["pairs"])
    assert e["derivatives"] and P >= 1 and e["n_active"].sum() > 0
    assert set(e["pairs"]) <= {("box", "ground"), ("ground", "box")}
    g = np.random.default_rng(5).normal(size=(P, 12))
    r = sim.contact_vjp(g)
    with pytest.raises(ValueError):
        sim.contact_vjp(np.zeros((P + 1, 12)))
    worst = 0.0
    for i in range(P):
        if not e["n_active"][i]:
            assert not np.any(r["q1"][i]) and not np.any(r["phi2"][i])
            continue
        gk, gd = g[i] @ e["K"][i], g[i] @ e["D"][i]
        sk = np.linalg.norm(g[i]) * np.abs(e["K"][i]).max()
        sd = np.linalg.norm(g[i]) * max(np.abs(e["D"][i]).max(), 1e-30)
        ek = np.abs(np.concatenate([r["q1"][i], r["q2"][i]]) - gk).max() / sk
        ed = np.abs(np.concatenate([r["phi1"][i], r["phi2"][i]]) - gd).max() / sd
        worst = max(worst, ek, ed)
    # the scale derivative: fm at scale 1 +- h, same state, no derivatives needed
    h = 1e-4
    fs = []
    for s in (1 + h, 1 - h):
        sim.set_contact_scale(s)
        sim.computeMatrices(False)
        fs.append(sim.last_contact_eval()["f"])
    sim.set_contact_scale(1.0)
    dfs = (fs[0] - fs[1]) / (2 * h)
    lhs = np.einsum("pi,pi->p", g, dfs)
    rhs = KN * r["kn"] + KT * r["kt"]
    scale = np.abs(g).max() * np.abs(dfs).max()
    print(f"  {P} pairs, {int(e['n_active'].sum())} active points: VJP vs g K / g D worst {worst:.3g}; "
          f"kn, kt vs the contact-scale difference {np.abs(lhs - rhs).max() / scale:.3g}", flush=True)
    assert worst <= 1e-10
    assert np.abs(lhs - rhs).max() <= 1e-6 * scale

"""The simulator subset on the host, without a GPU: the scene XML reader, OBJ reading and
mass properties (BodyMeshObj.cpp:47-177), the free3d-exp joint's kinematics, and the
reduced rigid-body dynamics and BDF1 steps (Simulation.cpp:379-506, 999-1098, 1248-1270)
for scenes of mesh bodies, which have no contact forces. Checked against closed forms,
finite differences, the Newton-Euler equations and the reference integrator."""
import os

import numpy as np
import pytest
from scipy.linalg import expm

import scene_meshes as SM
from atac import _redmax as R
from sim_reference import RefScene

OPTION = '<option integrator="BDF1" timestep="1e-3" gravity="0. 0. -9.8"/>'


def link(name, joint, body_type="mesh", filename="box.obj", jpos="0 0 0", jquat="1 0 0 0", bpos="0 0 0",
         bquat="1 0 0 0", extra="", damping="0"):
    return f'''<robot><link name="{name}">
      <joint name="{name}" type="{joint}" axis="0. 0. 0." pos="{jpos}" quat="{jquat}" frame="WORLD" damping="{damping}"/>
      <body name="{name}" type="{body_type}" filename="{filename}" pos="{bpos}" quat="{bquat}" scale="1 1 1"
            transform_type="OBJ_TO_JOINT" density="2" {extra}/></link></robot>'''


def scene(links, contacts="", option=OPTION, default=""):
    return f'<redmax model="test">{option}{default}{"".join(links)}<contact>{contacts}</contact></redmax>'


@pytest.fixture(scope="module")
def assets(tmp_path_factory):
    d = tmp_path_factory.mktemp("assets")
    V, F = SM.box_mesh([2.0, 1.0, 0.5], 3, center=(0.3, -0.2, 0.1))
    SM.write_obj(os.path.join(d, "box.obj"), V, F)
    return str(d)


# ------------------------------------------------------------------ scene XML

def test_xml_reader_errors(assets):
    for bad in ("<redmax><robot>", "<redmax a=1></redmax>", "<redmax></robot>", "<other/>", "<redmax a='x></redmax>"):
        with pytest.raises(RuntimeError):
            R.Simulation(bad, assets)
    # comments, declaration and entities are accepted
    s = R.Simulation('<?xml version="1.0"?><!-- c --><redmax model="a&amp;b"><option timestep="2e-3"/></redmax>', assets)
    assert s.ndof_r == 0 and s.timestep == np.float32(2e-3)


def test_scene_constructs_outside_the_subset_are_refused(assets):
    cases = [
        scene([link("a", "revolute")]),
        scene([link("a", "fixed", body_type="cuboid")]),
        scene([link("a", "fixed")], '<general_primitive_contact general_body="a" primitive_body="a"/>'),
        scene([link("a", "fixed")], '<general_SDF_contact general_body="a" SDF_body="nope"/>'),
        scene([link("a", "fixed"), link("b", "fixed")], '<general_SDF_contact general_body="a" SDF_body="b"/>'),
        '<redmax><ground pos="0 0 0" normal="0 0 1"/></redmax>',
        f'<redmax><robot><link name="a"><joint name="a" type="fixed" pos="0 0 0" quat="1 0 0 0"/>'
        f'<body name="a" type="mesh" filename="box.obj" pos="0 0 0" quat="1 0 0 0"/><link name="b"/></link></robot></redmax>',
    ]
    for xml in cases:
        with pytest.raises(RuntimeError):
            R.Simulation(xml, assets)


def test_options_defaults_and_float_attributes(assets):
    s = R.Simulation(scene([link("a", "translational")], option=""), assets)
    assert s.timestep == 0.02 and s.integrator == "BDF1"                   # Simulation.h:90
    np.testing.assert_array_equal(s.gravity.ravel(), [0.0, 0.0, -980.0])
    s = R.Simulation(scene([link("a", "translational")]), assets)
    assert s.timestep == float(np.float32(1e-3))                         # pugixml as_float
    np.testing.assert_array_equal(s.gravity.ravel(), [0.0, 0.0, -9.8])   # str_to_eigen: doubles


# ------------------------------------------------------------------ mesh and mass properties

def test_obj_reader(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# quad, relative indices, v/vt/vn tokens, a second shape\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                 "f 1/1/1 2/1/1 3/1/1 4/1/1\nf -4 -2 -1\no second\nf 1 2 3\n")
    V, F = R.load_obj(str(p), [2.0, 3.0, 1.0])
    np.testing.assert_array_equal(V, [[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0]])
    np.testing.assert_array_equal(F.T, [[0, 1, 2], [0, 2, 3], [0, 2, 3]])   # fan; the second shape dropped
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(RuntimeError):
        R.load_obj(str(bad))
    with pytest.raises(RuntimeError):
        R.load_obj(str(tmp_path / "missing.obj"))


def test_mass_properties_closed_forms():
    a, b, c, rho = 2.0, 1.0, 0.5, 3.0
    V, F = SM.box_mesh([a, b, c], 2, center=(1.0, -2.0, 0.5))
    # rotate the box so the principal axes are not the file's axes
    Rz = expm(np.array([[0, -0.7, 0.2], [0.7, 0, -0.4], [-0.2, 0.4, 0]]))
    V = V @ Rz.T
    p = R.process_mesh(V, F, rho)
    m = a * b * c * rho
    assert abs(p["volume"] - a * b * c) < 1e-12 and abs(p["mass"] - m) < 1e-12
    np.testing.assert_allclose(p["com"].ravel(), Rz @ [1.0, -2.0, 0.5], atol=1e-12)
    moments = np.sort(m / 12 * np.array([b * b + c * c, a * a + c * c, a * a + b * b]))
    np.testing.assert_allclose(p["inertia"].ravel(), np.concatenate([moments, [m] * 3]), rtol=1e-12)
    E = p["E_oi"]
    Rb = E[:3, :3]
    np.testing.assert_allclose(Rb.T @ Rb, np.eye(3), atol=1e-12)
    assert np.linalg.det(Rb) > 0
    # the principal axes diagonalise the inertia tensor
    Iw = p["I_unit_mass"] * m
    np.testing.assert_allclose(Rb.T @ Iw @ Rb, np.diag(moments), atol=1e-12)
    # the mesh in the body frame is centred and axis-aligned with the box
    Vb = p["V_body"]
    np.testing.assert_allclose(Vb.mean(axis=0), 0, atol=1e-12)   # symmetric lattice: the vertex mean is the COM
    np.testing.assert_allclose(np.sort(Vb.max(axis=0) - Vb.min(axis=0)), np.sort([a, b, c]), rtol=1e-12)
    # a tetrahedron: volume 1/6, centre of mass at the vertex mean
    T = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    Ft = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    pt = R.process_mesh(T, Ft, 1.0)
    assert abs(pt["volume"] - 1 / 6) < 1e-15
    np.testing.assert_allclose(pt["com"].ravel(), [0.25] * 3, atol=1e-15)


# ------------------------------------------------------------------ free3d-exp kinematics

def _skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


@pytest.mark.parametrize("r", [[1e-9, -2e-9, 3e-9], [0.1, -0.2, 0.15], [0.3, 0.2, -0.3], [1.0, -2.0, 1.5],
                               [3.0, 1.0, -2.0]])
def test_so3_derivatives(r):
    r = np.array(r)
    d = R.so3_derivatives(r)
    Rr = expm(_skew(r))
    np.testing.assert_allclose(d["R"], Rr, atol=1e-14)
    h = 1e-6
    for j in range(3):
        e = np.eye(3)[j]
        fd = (expm(_skew(r + h * e)) - expm(_skew(r - h * e))) / (2 * h)
        np.testing.assert_allclose(d["dR"][j], fd, atol=1e-9)
        # Jr column j: the body angular velocity per unit rdot_j, vee(R^T dR/dr_j)
        W = Rr.T @ d["dR"][j]
        np.testing.assert_allclose(d["Jr"][:, j], [W[2, 1], W[0, 2], W[1, 0]], atol=1e-13)
        dp, dm = R.so3_derivatives(r + h * e), R.so3_derivatives(r - h * e)
        np.testing.assert_allclose(d["dJr"][j], (dp["Jr"] - dm["Jr"]) / (2 * h), atol=1e-9)
        for k in range(3):
            np.testing.assert_allclose(d["d2R"][k][j], (dp["dR"][k] - dm["dR"][k]) / (2 * h), atol=1e-9)
            np.testing.assert_allclose(d["d2Jr"][k][j], (dp["dJr"][k] - dm["dJr"][k]) / (2 * h), atol=1e-9)


# ------------------------------------------------------------------ dynamics

@pytest.fixture()
def free_box(assets):
    xml = scene([link("box", "free3d-exp", jpos="0.1 0.2 0.3", jquat="0.9 0.1 0.2 0.3", bpos="0.05 -0.1 0.2",
                      bquat="1 0.2 0 0.1")])
    s = R.Simulation(xml, assets)
    rng = np.random.default_rng(0)
    q = np.concatenate([rng.normal(0, 0.3, 3), rng.normal(0, 0.8, 3)])
    qd = rng.normal(0, 1.0, 6)
    ext = rng.normal(0, 1.0, 6)
    return s, q, qd, ext


def _vee(X):
    return np.array([X[2, 1], X[0, 2], X[1, 0], X[0, 3], X[1, 3], X[2, 3]])


def test_reduced_dynamics_match_newton_euler(free_box):
    s, q, qd, ext = free_box
    s.set_body_external_force("box", ext)

    def pose(qq):
        s.set_state(qq, np.zeros(6))
        return s.get_body_E0i("box")

    def twist(qq, qqd, dt=1e-6):
        return _vee(np.linalg.inv(pose(qq)) @ (pose(qq + dt * qqd) - pose(qq - dt * qqd)) / (2 * dt))

    s.set_state(q, qd)
    o = s.computeMatrices(False)
    M, f = o["M"], o["f"]
    Mm = s.get_body_inertia("box").ravel()
    phi = twist(q, qd)
    assert abs(0.5 * qd @ M @ qd - 0.5 * phi @ (Mm * phi)) < 1e-9 * (0.5 * phi @ (Mm * phi))
    # q'' = M^-1 f must give the body the Newton-Euler acceleration:
    # Mm phi' = ad(phi)^T Mm phi + F, F = [I (R^T tau); m R^T (f + m g)]
    qdd = np.linalg.solve(M, f)
    dt = 1e-4
    phid = (twist(q + dt * qd + 0.5 * dt * dt * qdd, qd + dt * qdd) -
            twist(q - dt * qd + 0.5 * dt * dt * qdd, qd - dt * qdd)) / (2 * dt)
    Rw = pose(q)[:3, :3]
    w, v = phi[:3], phi[3:]
    adT = np.concatenate([np.cross(Mm[:3] * w, w) + np.cross(Mm[3:] * v, v), np.cross(Mm[3:] * v, w)])
    F = np.concatenate([Mm[:3] * (Rw.T @ ext[:3]), Mm[3:] * (Rw.T @ ext[3:]) + Rw.T @ (Mm[3] * np.array([0, 0, -9.8]))])
    assert np.abs(Mm * phid - adT - F).max() < 1e-5 * np.abs(F).max()


def test_derivatives_match_finite_differences(free_box):
    s, q, qd, ext = free_box
    s.set_body_external_force("box", ext)
    s.set_state(q, qd)
    o = s.computeMatrices(True)
    h = 1e-6
    for k in range(6):
        e = np.eye(6)[k]
        s.set_state(q + h * e, qd)
        a = s.computeMatrices(False)
        s.set_state(q - h * e, qd)
        b = s.computeMatrices(False)
        np.testing.assert_allclose(o["K"][:, k], (a["f"] - b["f"]) / (2 * h), atol=1e-7 * np.abs(o["K"]).max())
        np.testing.assert_allclose(o["dM_dq"][k], (a["M"] - b["M"]) / (2 * h), atol=1e-7 * np.abs(o["M"]).max())
        s.set_state(q, qd + h * e)
        a = s.computeMatrices(False)
        s.set_state(q, qd - h * e)
        b = s.computeMatrices(False)
        np.testing.assert_allclose(o["D"][:, k], (a["f"] - b["f"]) / (2 * h), atol=1e-7 * np.abs(o["D"]).max())


def test_bdf1_steps_match_reference_integrator(free_box):
    s, q, qd, ext = free_box
    s.set_state_init(q, qd)
    s.reset()
    s.set_body_external_force("box", ext)
    ref = RefScene(s, [("box", "free3d-exp", 0.0)], [], [0, 0, -9.8], s.timestep)
    ref.bodies[0].f_ext = ext
    ref.set_E_pj0(s, q)
    s.forward(20)
    assert s.is_converged()
    qr, qdr = q.copy(), qd.copy()
    for qs in s.get_q_his()[1:]:
        qr, qdr, res = ref.step(qr, qdr)
        assert res < 1e-9
        np.testing.assert_allclose(qs, qr, atol=1e-9)


def test_reparam_keeps_rotation_and_angular_velocity(assets):
    """JointSphericalExp::reparam (:237-262): past |r| = 1.5 pi the step's result is moved to
    the chart r1 = (1 - 2 pi / |r|) r with the same rotation and angular velocity."""
    h = 1e-2
    s = R.Simulation(scene([link("box", "free3d-exp")], option=f'<option timestep="{h}" gravity="0 0 0"/>'), assets)
    h = s.timestep
    q0 = np.array([0, 0, 0, 4.6, 0.2, -0.1])      # |r| just below 1.5 pi = 4.712
    qd0 = np.array([0, 0, 0, 20.0, 1.0, 0.0])     # spinning past it within the step
    s.set_state_init(q0, qd0)
    s.reset()
    s.forward(1)
    q1, qd1 = s.get_q(), s.get_qdot()
    r1 = q1[3:]
    assert np.linalg.norm(r1) <= 1.5 * np.pi
    # the step's solution before the chart change: antiparallel, |r_un| = 2 pi - |r1|
    r_un = -r1 / np.linalg.norm(r1) * (2 * np.pi - np.linalg.norm(r1))
    assert np.linalg.norm(r_un) > 1.5 * np.pi
    rd_un = (r_un - q0[3:]) / h
    d1, du = R.so3_derivatives(r1), R.so3_derivatives(r_un)
    np.testing.assert_allclose(d1["R"], du["R"], atol=1e-12)
    np.testing.assert_allclose(d1["Jr"] @ qd1[3:], du["Jr"] @ rd_un, atol=1e-9 * np.linalg.norm(rd_un))


def test_translational_and_fixed_bodies_and_state_api(assets):
    xml = scene([link("a", "translational", jpos="0 0 1", damping="0.5"), link("b", "fixed", jpos="3 0 0")])
    s = R.Simulation(xml, assets)
    assert s.ndof_r == 3 and s.ndof_m == 12
    m = s.get_body_mass("a")
    s.reset()
    s.set_body_external_force("a", [0, 0, 0, 1.0, 0, 0])
    s.forward(10)
    # translational joint: M = m I, f = m g + m R^T... applied as m * (R^T f) rotated back: the
    # joint sees m * f_ext (the reference scales the external force by the mass) plus gravity
    h = s.timestep
    q, qd = np.zeros(3), np.zeros(3)
    for qs in s.get_q_his()[1:]:
        # BDF1 with linear damping: m (q1 - q0 - h qd0) = h^2 (m f_ext + m g - c qd1), qd1 = (q1 - q0)/h
        rhs = m * (q + h * qd) + h * h * (m * np.array([1.0, 0, -9.8])) + h * 0.5 * q
        q1 = rhs / (m + h * 0.5)
        qd, q = (q1 - q) / h, q1
        np.testing.assert_allclose(qs, q, atol=1e-10)
    assert np.allclose(s.get_joint_q("a"), q)
    s.zero_joint_qdot("a")
    assert np.all(s.get_joint_qdot("a") == 0)
    s.set_joint_q("a", [1.0, 2.0, 3.0])
    E = s.get_body_E0i("a")   # [I, (0, 0, 1) + q] E_ji
    np.testing.assert_allclose(E[:3, 3], np.array([1.0, 2.0, 4.0]) + s.get_body_Eji("a")[:3, 3], atol=1e-14)
    np.testing.assert_allclose(E @ s.get_body_Ei0("a"), np.eye(4), atol=1e-14)
    Vb, Vw = s.get_body_vertices("a"), s.get_body_vertices("a", world_frame=True)
    np.testing.assert_allclose(Vw, E[:3, :3] @ Vb + E[:3, 3:4], atol=1e-13)
    assert s.get_body_faces("a").shape[0] == 3
    with pytest.raises(RuntimeError):
        s.get_joint_q("missing")
    with pytest.raises(RuntimeError):   # forward needs reset (Simulation.cpp:918-921)
        R.Simulation(xml, assets).forward(1)


def test_candidates_step_as_forward_does(free_box):
    """forward_candidates: each candidate ends exactly where reset + forward would leave it
    (no contact here: mesh bodies), and the scene's own state is left as it was."""
    s, q, qd, ext = free_box
    rng = np.random.default_rng(5)
    K = 4
    qs = np.stack([q + rng.normal(0, 0.1, 6) for _ in range(K)])
    qds = np.stack([qd + rng.normal(0, 0.5, 6) for _ in range(K)])
    fs = rng.normal(0, 2.0, (K, 1, 6))
    s.set_state(q, qd)
    out = s.forward_candidates(qs, qds, fs, 7)
    np.testing.assert_array_equal(s.get_q(), q)
    np.testing.assert_array_equal(s.get_qdot(), qd)
    for c in range(K):
        s.set_state_init(qs[c], qds[c])
        s.reset()
        s.set_body_external_force("box", fs[c, 0])
        s.forward(7)
        np.testing.assert_array_equal(out["q"][c], s.get_q())
        np.testing.assert_array_equal(out["qdot"][c], s.get_qdot())
        assert out["converged"][c] == int(s.is_converged())

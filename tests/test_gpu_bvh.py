"""GPU parity of the BVH contact variant (BodyBVHObj / ForceGeneralBVHContact)
against the brute-force FP64 oracle (oracle_mesh_distance_normal, oracle_bvh_eval):
signed distances, pseudonormals and closest faces through the C ABI, and the
wrench, p6 column, Km and Dm of whole force evaluations."""
import numpy as np
import pytest

from atac import _lib as L
from atac.bvh import BVHMesh
from atac.workloads import box_mesh, icosphere, torus
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MESHES = {
    "cube": lambda: box_mesh(0.6, center=(0.05, -0.1, 0.02)),
    "sphere": lambda: icosphere(0.8, levels=3),
    "torus": lambda: torus(1.0, 0.35, 32, 16),
}


@pytest.fixture(scope="module", autouse=True)
def _device():
    from atac.batch import device_init
    device_init(0)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_distance_normal_matches_oracle(name):
    V, F = MESHES[name]()
    rng = np.random.default_rng(11)
    E2 = O.se3_exp(rng.normal(size=6) * 0.3)
    lo, hi = V.min(0) - 0.1, V.max(0) + 0.1
    xb = rng.uniform(lo, hi, size=(4000, 3))
    xw = xb @ E2[:3, :3].T + E2[:3, 3]
    m = BVHMesh(V, F)
    d, n, f = m.distance_normal(xw, E2, col_th=0.01)
    od, on, of = O.mesh_distance_normal(V, F, xw, E2=E2, col_th=0.01)
    print(name, "inside box", int((od != 10.0).sum()), "of", len(od), flush=True)
    np.testing.assert_allclose(d, od, rtol=0, atol=1e-12)
    out = od == 10.0
    assert np.all(f[out] == -1) and np.all(n[out] == 0)
    # a closest point on an edge or vertex shared by several triangles is a tie
    # between them, which the last bit of the FP64 distances (the GPU fuses
    # multiply-adds) can resolve either way; the pseudonormal of the shared
    # feature is the same for all of them
    same = f == of
    agree = np.all(np.abs(n - on) <= 1e-12, axis=1)
    print(name, "faces equal", same.mean(), "normals equal", agree.mean(), flush=True)
    assert agree[same].all()
    assert agree.mean() > 0.999, agree.mean()


def test_bad_meshes():
    V, F = box_mesh(1.0)
    with pytest.raises(L.AtacError):
        BVHMesh(V, np.array([[0, 1, 99]]))
    with pytest.raises(L.AtacError):
        BVHMesh(np.zeros((2, 3)), np.array([[0, 1, 1]]))
    bad = V.copy()
    bad[0, 0] = np.nan
    with pytest.raises(L.AtacError):
        BVHMesh(bad, F)


def _scene(name, seed):
    V, F = MESHES[name]()
    rng = np.random.default_rng(seed)
    P, _ = icosphere(0.35, levels=3)           # contact points of a small ball
    xi = np.concatenate([P, P * 0.9])
    E1 = O.se3_exp(np.concatenate([rng.normal(size=3) * 0.3, np.zeros(3)]))
    # put the ball across the mesh surface: centre near a surface point
    c = V[rng.integers(len(V))] * 1.02
    E1[:3, 3] = c
    E2 = np.eye(4)
    phi1, phi2 = rng.normal(size=6) * 0.2, rng.normal(size=6) * 0.1
    return V, F, xi, E1, phi1, E2, phi2


@pytest.mark.parametrize("name,seed", [("cube", 1), ("sphere", 2), ("torus", 3), ("torus", 4)])
def test_force_matches_oracle(name, seed):
    V, F, xi, E1, phi1, E2, phi2 = _scene(name, seed)
    args = dict(col_th=0.005, kn=3e4, kt=1e3, mu=0.6, damping=5.0, scale=1.3, contact_scale=0.9)
    m = BVHMesh(V, F)
    inst = O.make_instance(0, xi, E1, phi1, E2, phi2, args["col_th"], args["kn"], args["kt"], args["mu"],
                           args["damping"], args["scale"], args["contact_scale"])
    ref = O.bvh_eval(V, F, inst, flags=3)
    got = m.force(xi, E1, phi1, E2, phi2, deriv=True, p6=True, **args)
    print(name, seed, "contacts", got["n_active"], "of", len(xi), flush=True)
    assert got["n_active"] == ref["n_active"] and got["n_active"] > 10
    for key in ("fm", "dfm_dp6", "Km", "Dm"):
        a, b = np.asarray(got[key]), np.asarray(ref[key])
        scale = max(np.abs(b).max(), 1e-12)
        assert np.abs(a - b).max() <= 1e-9 * scale, (key, np.abs(a - b).max(), scale)
    plain = m.force(xi, E1, phi1, E2, phi2, **args)
    np.testing.assert_allclose(plain["fm"], got["fm"], rtol=0, atol=1e-9 * np.abs(got["fm"]).max())


def test_no_contact():
    V, F = MESHES["sphere"]()
    m = BVHMesh(V, F)
    xi = np.random.default_rng(0).normal(size=(100, 3))
    E1 = np.eye(4)
    E1[:3, 3] = [5.0, 0.0, 0.0]
    got = m.force(xi, E1, np.zeros(6), np.eye(4), np.zeros(6), 0.0, 1.0, 1.0, 0.5, 0.1, deriv=True, p6=True)
    assert got["n_active"] == 0
    assert np.all(got["fm"] == 0) and np.all(got["Km"] == 0) and np.all(got["Dm"] == 0)


def test_host_classes_scatter():
    """BodyBVHObj / ForceGeneralBVHContact (pybind): the global fm, Km, Dm and the
    contact-scale column hold the per-object result at each body's _index."""
    from atac import _redmax as R
    V, F, xi, E1, phi1, E2, phi2 = _scene("torus", 5)
    args = dict(col_th=0.004, kn=2e4, kt=5e2, mu=0.4, damping=3.0)
    a = R.Body("ball", xi, 6)
    b = R.BodyBVHObj("torus", V, F, args["col_th"], np.zeros((0, 3)), 18)
    a.E_0i, a.phi, b.E_0i, b.phi = E1, phi1, E2, phi2
    a.contact_scale = 0.8
    a.p6_column = 2
    f = R.ForceGeneralBVHContact(a, b, args["kn"], args["kt"], args["mu"], args["damping"])
    fm, Km, Dm, dp = f.computeForceWithDesignDerivative(30, 4)
    inst = O.make_instance(0, xi, E1, phi1, E2, phi2, args["col_th"], args["kn"], args["kt"], args["mu"],
                           args["damping"], 1.0, 0.8)
    ref = O.bvh_eval(V, F, inst, flags=3)
    assert f.last_active_count() == ref["n_active"] > 0
    idx = np.r_[6:12, 18:24]
    s = np.abs(ref["fm"]).max()
    np.testing.assert_allclose(fm[idx], ref["fm"], atol=1e-9 * s)
    np.testing.assert_allclose(Km[np.ix_(idx, idx)], ref["Km"], atol=1e-9 * np.abs(ref["Km"]).max())
    np.testing.assert_allclose(Dm[np.ix_(idx, idx)], ref["Dm"], atol=1e-9 * np.abs(ref["Dm"]).max())
    np.testing.assert_allclose(dp[idx, 2], ref["dfm_dp6"], atol=1e-9 * np.abs(ref["dfm_dp6"]).max())
    rest = np.setdiff1d(np.arange(30), idx)
    assert np.all(fm[rest] == 0)
    # distance queries through the body
    xb = np.random.default_rng(1).uniform(V.min(0), V.max(0), size=(50, 3))
    d, n = b.distance_normal_parallel(xb @ E2[:3, :3].T + E2[:3, 3])
    od, on, of = O.mesh_distance_normal(V, F, xb @ E2[:3, :3].T + E2[:3, 3], E2=E2, col_th=args["col_th"])
    np.testing.assert_allclose(d, od, atol=1e-12)


def test_empty_inputs():
    """No query points: empty distance arrays; a force evaluation over no points is zero."""
    V, F = MESHES["cube"]()
    m = BVHMesh(V, F)
    d, n, f = m.distance_normal(np.zeros((0, 3)))
    assert d.shape == (0,) and n.shape == (0, 3) and f.shape == (0,)
    got = m.force(np.zeros((0, 3)), np.eye(4), np.zeros(6), np.eye(4), np.zeros(6), 0.01, 1e4, 1e3, 0.5, 1.0,
                  deriv=True, p6=True)
    assert got["n_active"] == 0
    for key in ("fm", "dfm_dp6", "Km", "Dm"):
        assert np.all(np.asarray(got[key]) == 0), key


def test_degenerate_triangles():
    """A mesh holding zero-area triangles (a collinear sliver floating beside a
    cube face, and a triangle with a repeated vertex) gives the oracle's
    distances, normals and faces: their face, edge and vertex normals are zero,
    so the points nearest to them get a positive sign."""
    V, F = box_mesh(0.5)
    V = np.vstack([V, [[0.6, -0.2, 0.1], [0.6, 0.0, 0.1], [0.6, 0.2, 0.1]]])
    nv = len(V)
    F = np.vstack([F, [[nv - 3, nv - 2, nv - 1], [nv - 3, nv - 3, nv - 1]]])
    rng = np.random.default_rng(5)
    xw = rng.uniform(-0.5, 0.5, size=(3000, 3))
    xw[:1000] = rng.uniform([0.5, -0.3, 0.0], [0.6, 0.3, 0.2], size=(1000, 3))   # around the slivers
    m = BVHMesh(V, F)
    d, n, f = m.distance_normal(xw, col_th=0.003)
    od, on, of = O.mesh_distance_normal(V, F, xw, col_th=0.003)
    np.testing.assert_allclose(d, od, rtol=0, atol=1e-12)
    same = f == of
    agree = np.all(np.abs(n - on) <= 1e-12, axis=1)
    print("faces equal", same.mean(), "normals equal", agree.mean(), "on slivers", int((of >= 12).sum()), flush=True)
    assert (of >= 12).sum() > 20
    assert agree[same].all()
    assert agree.mean() > 0.999


def test_force_many_workgroups():
    """A force evaluation over several 2,048-point workgroups (their partial rows
    added in workgroup order) equals the oracle."""
    V, F = MESHES["torus"]()
    rng = np.random.default_rng(8)
    P, _ = icosphere(0.35, levels=4)
    xi = np.concatenate([P, P * 0.95, P * 0.9])           # 7,686 points: 4 workgroups
    E1 = O.se3_exp(np.concatenate([rng.normal(size=3) * 0.3, np.zeros(3)]))
    E1[:3, 3] = V[rng.integers(len(V))] * 1.02
    E2 = np.eye(4)
    phi1, phi2 = rng.normal(size=6) * 0.2, rng.normal(size=6) * 0.1
    args = dict(col_th=0.005, kn=3e4, kt=1e3, mu=0.6, damping=5.0, scale=1.1, contact_scale=0.9)
    inst = O.make_instance(0, xi, E1, phi1, E2, phi2, args["col_th"], args["kn"], args["kt"], args["mu"],
                           args["damping"], args["scale"], args["contact_scale"])
    ref = O.bvh_eval(V, F, inst, flags=3)
    got = BVHMesh(V, F).force(xi, E1, phi1, E2, phi2, deriv=True, p6=True, **args)
    print("points", len(xi), "contacts", got["n_active"], flush=True)
    assert got["n_active"] == ref["n_active"] > 50
    for key in ("fm", "dfm_dp6", "Km", "Dm"):
        a, b = np.asarray(got[key]), np.asarray(ref[key])
        assert np.abs(a - b).max() <= 1e-9 * max(np.abs(b).max(), 1e-12), key

"""CPU tests of the BVH-variant oracle (brute-force signed distance with
pseudonormals, oracle/atac_oracle.cpp mesh_query), pinned by what the geometry
fixes rather than by retyping its formulas:
  * the cube: closed-form signed distance and normals of an axis-aligned box;
  * the sign: the generalized winding number (sum of triangle solid angles, van
    Oosterom & Strackee) of closed meshes, inside <=> |w| ~ 1;
  * the magnitude: an independent numpy point-triangle distance (projection onto
    the plane when it falls inside the triangle, else the nearest of the three
    edge segments);
  * the normals: unit length, face normals orthogonal to their triangle, cube
    vertex normals (+-1, +-1, +-1)/sqrt(3), coplanar-edge normals = face normal.
libigl itself (the reference's dependency, not vendored) is not available, so
these properties are its specification here."""
import numpy as np
import pytest

from atac.workloads import box_mesh, icosphere, torus
from oracle import oracle as O


def box_sdf(p, a):
    q = np.abs(p) - a
    out = np.linalg.norm(np.maximum(q, 0.0), axis=1)
    ins = np.minimum(np.max(q, axis=1), 0.0)
    return out + ins


def winding(V, F, p):
    a = V[F[:, 0]][None] - p[:, None]
    b = V[F[:, 1]][None] - p[:, None]
    c = V[F[:, 2]][None] - p[:, None]
    la, lb, lc = (np.linalg.norm(x, axis=2) for x in (a, b, c))
    det = np.einsum("pfi,pfi->pf", a, np.cross(b, c))
    den = la * lb * lc + np.einsum("pfi,pfi->pf", a, b) * lc + np.einsum("pfi,pfi->pf", a, c) * lb + \
        np.einsum("pfi,pfi->pf", b, c) * la
    return np.sum(2.0 * np.arctan2(det, den), axis=1) / (4.0 * np.pi)


def seg_dist(p, a, b):
    ab = b - a
    t = np.clip(np.einsum("pfi,pfi->pf", p - a, ab) / np.maximum(np.einsum("pfi,pfi->pf", ab, ab), 1e-300), 0, 1)
    return np.linalg.norm(p - (a + t[..., None] * ab), axis=2)


def tri_dist(V, F, p):
    P = p[:, None, :]
    a, b, c = V[F[:, 0]][None], V[F[:, 1]][None], V[F[:, 2]][None]
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n, axis=2, keepdims=True)
    h = np.einsum("pfi,pfi->pf", P - a, n)
    q = P - h[..., None] * n
    inside = np.ones(h.shape, dtype=bool)
    for u, v in ((a, b), (b, c), (c, a)):
        inside &= np.einsum("pfi,pfi->pf", np.cross(v - u, q - u), n) >= 0
    d_edge = np.minimum(np.minimum(seg_dist(P, a, b), seg_dist(P, b, c)), seg_dist(P, c, a))
    return np.min(np.where(inside, np.abs(h), d_edge), axis=1)


def test_cube_closed_form():
    V, F = box_mesh(0.7, center=(0.1, -0.2, 0.3))
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.7, 0.7, size=(3000, 3)) + np.array([0.1, -0.2, 0.3])
    d, n, f = O.mesh_distance_normal(V, F, p)
    ref = box_sdf(p - np.array([0.1, -0.2, 0.3]), 0.7)
    assert np.all(d <= 1e-15)
    np.testing.assert_allclose(d, ref, atol=1e-13)
    # interior points: the normal is the axis of the nearest face
    q = p - np.array([0.1, -0.2, 0.3])
    ax = np.argmax(np.abs(q), axis=1)
    clear = np.sort(np.abs(q), axis=1)[:, 2] - np.sort(np.abs(q), axis=1)[:, 1] > 1e-9
    expect = np.zeros_like(q)
    expect[np.arange(len(q)), ax] = np.sign(q[np.arange(len(q)), ax])
    np.testing.assert_allclose(n[clear], expect[clear], atol=1e-14)
    # outside the vertices' bounding box: d = 10 + 0, normal 0, face -1
    d2, n2, f2 = O.mesh_distance_normal(V, F, np.array([[2.0, 0.0, 0.0]]), col_th=0.01)
    assert d2[0] == 10.0 and np.all(n2 == 0) and f2[0] == -1


def test_cube_normals():
    V, F = box_mesh(1.0)
    FN, FE, EN, VN = O.mesh_normals(V, F)
    np.testing.assert_allclose(np.abs(VN), 1 / np.sqrt(3), atol=1e-15)
    np.testing.assert_allclose(np.sign(VN), np.sign(V), atol=0)
    for f in range(len(F)):
        e1, e2 = V[F[f, 1]] - V[F[f, 0]], V[F[f, 2]] - V[F[f, 0]]
        assert abs(np.dot(FN[f], e1)) < 1e-15 and abs(np.dot(FN[f], e2)) < 1e-15
        assert np.dot(FN[f], V[F[f, 0]]) > 0   # outward
    # an edge shared by two coplanar triangles has their normal; a cube edge the
    # normalised sum of the two face normals
    owners = {}
    for f in range(len(F)):
        for c in range(3):
            owners.setdefault(FE[f, c], []).append(f)
    for e, fs in owners.items():
        assert len(fs) == 2
        s = FN[fs[0]] + FN[fs[1]]
        np.testing.assert_allclose(EN[e], s / np.linalg.norm(s), atol=1e-15)


@pytest.mark.parametrize("name", ["sphere", "torus"])
def test_sign_matches_winding_number(name):
    V, F = icosphere(1.0, levels=2) if name == "sphere" else torus(1.0, 0.4, 20, 10)
    rng = np.random.default_rng(1)
    lo, hi = V.min(0), V.max(0)
    p = rng.uniform(lo, hi, size=(1500, 3))
    d, n, f = O.mesh_distance_normal(V, F, p)
    w = winding(V, F, p)
    keep = np.abs(d) > 1e-9
    assert keep.sum() > 1400
    np.testing.assert_array_equal(d[keep] < 0, np.abs(w[keep]) > 0.5)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-14)


@pytest.mark.parametrize("name", ["sphere", "torus"])
def test_magnitude_matches_independent_distance(name):
    V, F = icosphere(1.0, levels=2) if name == "sphere" else torus(1.0, 0.4, 16, 8)
    rng = np.random.default_rng(2)
    lo, hi = V.min(0), V.max(0)
    p = rng.uniform(lo, hi, size=(600, 3))
    d, n, f = O.mesh_distance_normal(V, F, p)
    np.testing.assert_allclose(np.abs(d), tri_dist(V, F, p), rtol=1e-12, atol=1e-13)
    # the reported face holds the closest point
    own = np.array([tri_dist(V, F[f[i]][None], p[i][None])[0] for i in range(len(p))])
    np.testing.assert_allclose(own, np.abs(d), rtol=1e-12, atol=1e-13)


def test_pose_and_col_th():
    V, F = icosphere(0.5, levels=2)
    rng = np.random.default_rng(3)
    q = rng.normal(size=6) * 0.5
    E2 = O.se3_exp(q)
    x_body = rng.uniform(-0.5, 0.5, size=(200, 3))
    xw = x_body @ E2[:3, :3].T + E2[:3, 3]
    d_body, n_body, f_body = O.mesh_distance_normal(V, F, x_body)
    d_w, n_w, f_w = O.mesh_distance_normal(V, F, xw, E2=E2, col_th=0.02)
    ins = d_body != 10.0
    np.testing.assert_allclose(d_w[ins], d_body[ins] + 0.02, atol=1e-13)
    np.testing.assert_array_equal(f_w, f_body)


def test_degenerate_triangle_is_its_edges():
    """A zero-area triangle (three collinear vertices, or a repeated vertex) is
    taken as its edges: the distance to a collinear sliver is the distance to its
    segment, and its normals are zero, so that distance is positive."""
    V, F = box_mesh(0.5)
    V = np.vstack([V, [[0.6, -0.2, 0.1], [0.6, 0.0, 0.1], [0.6, 0.2, 0.1]]])
    nv = len(V)
    F = np.vstack([F, [[nv - 3, nv - 2, nv - 1], [nv - 3, nv - 3, nv - 1]]])
    FN, FE, EN, VN = O.mesh_normals(V, F)
    assert np.all(FN[12:] == 0) and np.all(VN[nv - 3:] == 0)
    rng = np.random.default_rng(6)
    p = rng.uniform([0.5, -0.3, 0.0], [0.6, 0.3, 0.2], size=(2000, 3))
    d, n, f = O.mesh_distance_normal(V, F, p)
    on = f == 12
    assert on.sum() > 100
    t = np.clip((p[on, 1] + 0.2) / 0.4, 0.0, 1.0)
    q = np.stack([np.full(on.sum(), 0.6), -0.2 + 0.4 * t, np.full(on.sum(), 0.1)], axis=1)
    np.testing.assert_allclose(d[on], np.linalg.norm(p[on] - q, axis=1), rtol=1e-13, atol=1e-15)
    assert np.all(n[on] == 0)
    # the others are the cube's: closed form of the box
    np.testing.assert_allclose(d[~on], box_sdf(p[~on], 0.5), atol=1e-13)

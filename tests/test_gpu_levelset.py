"""The GPU mesh -> SDF builder (atac_sdf_from_mesh, the level-set construction of
sdfgen::make_level_set3) against the CPU restatement: the same FP32 value at every
node, bit for bit, for convex and non-convex meshes, exact bands 0-2, uneven and
degenerate grid sizes; the grid placement of precompute_SDF; and SDFMesh (Mesh.cpp
SDFMesh: construction, distance, min_distance) against the oracle's query_SDF."""
import numpy as np
import pytest

import meshes as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def _box_with_slivers():
    """A box plus two zero-area triangles beside it: three collinear vertices, and
    a triangle with a repeated vertex."""
    V, F = M.box(1.0)
    V = np.vstack([V, [[1.2, -0.4, 0.05], [1.2, 0.1, 0.05], [1.2, 0.4, 0.05]]])
    n = len(V)
    return V, np.vstack([F, [[n - 3, n - 2, n - 1], [n - 3, n - 3, n - 1]]])


CASES = [
    ("box", M.box(1.0), (-1.53, -1.47, -1.51), (0.1, 0.1, 0.1), (31, 30, 32), 1),
    ("box_slivers", _box_with_slivers(), (-1.53, -1.47, -1.51), (0.1, 0.1, 0.1), (31, 30, 32), 1),
    ("box_band0", M.box(0.8, (0.1, -0.05, 0.02)), (-1.0, -1.0, -1.0), (0.07, 0.09, 0.11), (29, 23, 19), 0),
    ("sphere", M.icosphere(1.0, 3), (-1.31, -1.29, -1.3), (0.05, 0.05, 0.05), (53, 52, 53), 1),
    ("torus_band2", M.torus(1.0, 0.35, 32, 16), (-1.5, -1.5, -0.5), (0.04, 0.045, 0.05), (76, 68, 21), 2),
    ("torus_coarse", M.torus(1.0, 0.35, 24, 12), (-1.6, -1.6, -0.6), (0.3, 0.3, 0.3), (12, 12, 5), 1),
    ("flat", M.box(1.0), (-1.2, -1.2, -1.2), (0.1, 0.1, 0.7), (25, 25, 1), 1),
    ("line", M.box(1.0), (-1.2, -0.3, 0.1), (0.1, 0.7, 0.7), (25, 1, 1), 1),
    ("single", M.box(1.0), (0.2, 0.1, -0.3), (0.5, 0.5, 0.5), (1, 1, 1), 1),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_level_set_matches_oracle_bitwise(gpu, case):
    from atac.sdfmesh import level_set
    _, (V, F), origin, d, n, band = case
    ref = O.make_level_set3(V.astype(np.float32), F, origin, d, n, band)
    got = level_set(V, F, origin, d, n, band)
    assert got.shape == ref.shape
    bad = np.argwhere(np.asarray(got).view(np.uint32) != np.asarray(ref).view(np.uint32))
    assert len(bad) == 0, f"{len(bad)} nodes differ, first {bad[:3].tolist()}: {got[tuple(bad[0])]} vs {ref[tuple(bad[0])]}"


def test_mesh_grid_placement_matches_oracle(gpu):
    from atac.sdfmesh import mesh_grid
    for V, _ in (M.torus(1.0, 0.35), M.icosphere(0.7, 2, (0.3, -0.2, 5.0)), M.box(2.0)):
        for dx, ms in ((0.05, 20), (0.5, 20), (0.02, 64)):
            a = mesh_grid(V, dx, ms)
            b = O.sdf_mesh_grid(V, dx, ms)
            assert _same_bits(a[0], b[0]) and _same_bits(a[1], b[1])
            assert np.array_equal(a[2], b[2]) and a[3] == b[3]


def test_sdfmesh_build_and_distance(gpu, tmp_path):
    from atac.sdfmesh import SDFMesh
    V, F = M.torus(1.0, 0.35, 32, 16)
    save = str(tmp_path / "torus.sdf")
    m = SDFMesh(V.T, F.T, 0.05, save_path=save)     # Eigen-style 3 x n inputs
    mn, mx, d, n = O.sdf_mesh_grid(V, 0.05, SDFMesh.MIN_SIZE)
    ref = O.make_level_set3(V.astype(np.float32), F, mn, d.astype(np.float32), n)
    assert _same_bits(m.values, ref)
    rng = np.random.default_rng(3)
    X = rng.uniform(mn - 0.2, mx + 0.2, size=(4000, 3))
    got = m.distance(X.T)
    G = O.GridRef(m.values, m.min_box, m.max_box, m.dxyz)
    want = O.query_sdf(G, X)
    inside = np.all((X >= m.min_box) & (X <= m.max_box), axis=1)
    assert inside.sum() > 1000 and (~inside).sum() > 100
    assert np.all(got[~inside] == np.float32(np.finfo(np.float32).max))
    scale = np.abs(m.values).max()
    assert np.max(np.abs(got[inside] - want[inside])) <= 1e-6 * scale
    assert m.min_distance(X.T) == got.min()
    # the saved text grid reads back rounded to 6 significant digits
    m2 = SDFMesh(V.T, F.T, 0.05, load_path=save)
    assert np.allclose(m2.values, m.values, rtol=1e-5, atol=1e-6)
    assert np.allclose(m2.min_box, m.min_box, rtol=1e-5)


def test_bad_mesh_inputs(gpu):
    import atac
    from atac.sdfmesh import level_set
    V, F = M.box(1.0)
    bad = F.copy()
    bad[3, 1] = 99
    with pytest.raises(atac.AtacError):
        level_set(V, bad, (-1.2, -1.2, -1.2), (0.1, 0.1, 0.1), (25, 25, 25))
    with pytest.raises(atac.AtacError):
        level_set(V, F, (-1.2, -1.2, -1.2), (0.1, 0.0, 0.1), (25, 25, 25))


def test_body_sdf_from_mesh(gpu):
    """BodySDFObj built from a mesh (BodySDFObj.cpp:38-77, spacing min(dx, extent / res))
    holds the oracle's grid, and its distance query is the oracle's."""
    from atac import _redmax as R
    V, F = M.icosphere(0.8, 2, (0.1, 0.2, -0.1))
    res = 30
    body = R.BodySDFObj.from_mesh("sphere", V, F, 0.05, res, 0.01, np.zeros((1, 3)))
    mn, mx, d, n = O.sdf_mesh_grid(V, 0.05, res)
    ref = O.make_level_set3(V.astype(np.float32), F, mn, d.astype(np.float32), n)
    assert tuple(body.shape) == n
    assert np.array_equal(body.values(), ref.astype(np.float64))
    assert np.array_equal(np.ravel(body.min_box), mn.astype(np.float64))
    assert np.array_equal(np.ravel(body.max_box), mx.astype(np.float64))
    assert np.array_equal(np.ravel(body.dxyz), d)
    G = O.GridRef(ref, mn.astype(np.float64), mx.astype(np.float64), d)
    rng = np.random.default_rng(4)
    for x in rng.uniform(mn, mx, size=(50, 3)):
        want = O.distance(G, np.eye(4), 0.01, x[None])[0]
        assert abs(body.distance(x) - want) <= 1e-6 * np.abs(ref).max()

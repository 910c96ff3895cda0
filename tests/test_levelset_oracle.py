"""The CPU restatement of the mesh -> SDF construction (sdfgen::make_level_set3,
makelevelset3.cpp:163-230) pinned against closed forms: exact distances near the
surface, signs everywhere, distances never below the true distance (every value is
the distance to some triangle), and the grid placement of precompute_SDF
(BodySDFObj.cpp:38-77, Mesh.cpp:54-95)."""
import numpy as np
import pytest

import meshes as M
from oracle import oracle as O


def _nodes(origin, d, n):
    ni, nj, nk = n
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    o = np.asarray(origin, dtype=np.float32)
    dd = np.asarray(d, dtype=np.float32)
    return np.stack([(i.astype(np.float32) * dd[0] + o[0]), (j.astype(np.float32) * dd[1] + o[1]),
                     (k.astype(np.float32) * dd[2] + o[2])], axis=-1).astype(np.float64)


@pytest.mark.parametrize("band", [0, 1, 2])
def test_box_exact_band_and_signs(band):
    V, F = M.box(1.0)
    origin, d, n = (-1.53, -1.47, -1.51), (0.1, 0.1, 0.1), (31, 30, 32)
    phi, cl = O.make_level_set3(V, F, origin, d, n, band, with_closest=True)
    p = _nodes(origin, d, n)
    ex = M.box_exact_sdf(p)
    near = np.abs(ex) < 0.1 * (band + 0.5)
    assert near.sum() > 500
    assert np.max(np.abs(phi[near] - ex[near])) < 2e-6
    assert np.all(np.sign(phi[np.abs(ex) > 1e-4]) == np.sign(ex[np.abs(ex) > 1e-4]))
    # every value is the distance to some triangle: never below the true distance
    assert np.all(np.abs(phi) >= np.abs(ex) - 2e-6)
    # a convex box: the sweeps carry the nearest face everywhere
    assert np.max(np.abs(phi - ex)) < 2e-6
    assert np.all(cl >= 0)


def test_sphere_signs_and_distance():
    V, F = M.icosphere(1.0, levels=3)
    origin, d, n = (-1.31, -1.29, -1.3), (0.05, 0.05, 0.05), (53, 52, 53)
    phi = O.make_level_set3(V, F, origin, d, n)
    p = _nodes(origin, d, n)
    rad = np.linalg.norm(p, axis=-1)
    # the polyhedron lies between its inradius and the unit sphere
    inr = 0.98
    assert np.all(phi[rad < inr - 1e-3] < 0) and np.all(phi[rad > 1.0 + 1e-3] > 0)
    far = rad > 1.05
    assert np.max(np.abs(phi[far] - (rad[far] - 1.0))) < 0.03


def test_torus_signs():
    V, F = M.torus(1.0, 0.35, 32, 16)
    origin, d, n = (-1.5, -1.5, -0.5), (0.04, 0.045, 0.05), (76, 68, 21)
    phi = O.make_level_set3(V, F, origin, d, n)
    p = _nodes(origin, d, n)
    ex = M.torus_exact_sdf(p)
    clear = np.abs(ex) > 0.03     # the polygonal tube differs from the round one by < 0.02
    assert np.all(np.sign(phi[clear]) == np.sign(ex[clear]))
    assert np.all(np.abs(phi) >= np.abs(ex) - 0.03)


def test_tiny_grids():
    V, F = M.box(1.0)
    for n in ((1, 1, 1), (2, 1, 3), (1, 4, 1)):
        phi = O.make_level_set3(V, F, (-1.2, -1.2, -1.2), (0.7, 0.7, 0.7), n)
        assert phi.shape == (n[2], n[1], n[0]) and np.all(np.isfinite(phi))


def test_grid_placement():
    V, _ = M.torus(1.0, 0.35)
    for dx, ms in ((0.05, 20), (0.5, 20), (0.01, 64)):
        mn, mx, d, n = O.sdf_mesh_grid(V, dx, ms)
        ext = V.max(0) - V.min(0)
        assert np.allclose(d, np.minimum(dx, ext / ms))
        # one cell of padding around the FP32 bounds, node counts rounding the box extent
        assert np.allclose(mn, V.min(0).astype(np.float32) - d.astype(np.float32), atol=1e-6)
        assert np.allclose(mx, V.max(0).astype(np.float32) + d.astype(np.float32), atol=1e-6)
        cells = (mx.astype(np.float64) - mn) / d
        assert np.all(np.abs((np.array(n) - 1) - cells) <= 0.5 + 1e-4)
